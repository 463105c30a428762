#!/usr/bin/env python3
"""Throughput of the polar-grid PM step on one MI355X.

Workload: 256 x 512 log-polar grid, N = 10M particles of the seeded
exponential disc (Q = 1.5, rigid flat-rotation halo), harmonics m <= 16,
three time-step zones (nstep 1, 2, 4; dt = 0.025, 0.05, 0.1), synthetic
initial conditions (galaxy_amd/configs.py C2).  One step = galaxy's loop
body: findf (zone combination, azimuthal transform, radial Green's
contraction, synthesis, polcat) then step (interpolation, halo force,
leap-frog, zone changes, mass assignment), all on the device; particles are
re-sorted by cell every --sort-interval steps inside the timed region.

value = particle pushes per second of the mixed form (FP32 particles and
grids, FP32 contraction): one push = one stpgrp application; with zones,
zone-k particles are pushed every nstep(k) steps.  The plain FP64 form runs
the same steps from the same particles, and the mixed form with compensated
position updates beside it; the JSON line carries their rates and the
ratio.  An accuracy leg outside the timed region, independent of --steps,
runs both forms over the workload's own 500 steps from the same particles:
the mixed solve's relative force error against the FP64 solve at the
start, each form's energy and Lz drift over the 500 steps and their
difference, and the m = 2 amplitude of both end states.

roofline: the push kernel, algorithmic bytes = pushes x 32 B (x, y, vx, vy
read and written in FP32) over its summed HIP-event time on the library's
stream, against 8 TB/s; traffic = measured HBM bytes per launch (per-push
bytes from rocprofv3 counters times the pushes per launch); step_frac = the
whole step's algorithmic bytes (pushes x 32 B, the two field types' FP32
tables, the zone mass sums, the field written and read) over the wall-clock
step time.  cpu_baseline: the reference's own galaxy program (built from
its Fortran sources into oracle/_ref; serial, one core) on a bounded sample
of the same workload (1e6 particles, 20 steps), or the oracle's FP32
restatement where that build is absent; extra.cpu_port_parallel: the
oracle's particle loop on every host thread at the full 1e7 particles.

--workload C4 measures the ensemble of 64 independent discs on 128 x 256
grids (N_i log-uniform in [1e5, 4e6], Q_i in [1, 2], 300 steps) advanced by
one launch per kernel, the timed steps replayed as HIP graphs: pushes over
all members per second, push-kernel roofline at 32 B per push, the oracle
on the largest member's disc as the CPU baseline.

--workload C3 measures the 3-D polar grid instead (200 x 256 x 125 grid,
N = 5e7, m <= 8, sech^2 vertical profile, 200 steps by default, re-sorted
every 8) and prints
its own line: push-kernel roofline at 48 B per push ((x, y, z, vx) and
(vy, vz) read and written), the p3d oracle as the CPU baseline.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("particle-steps/s @10M (mixed vs FP64 GPU, m=2 err<1e-4); "
          "% HBM roofline")
HBM_PEAK_GBS = 8000.0
# HBM bytes per push of push_kernel<float> on this workload: rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE of the all-zones launch (10M pushes), each scaled
# by what the same counter reads for a 1 GiB copy kernel (FETCH_SIZE x 2,
# WRITE_SIZE x 1); profiles/push_kernel_p2d-256x512-n1e7-3zones_traffic.txt.
PUSH_HBM_READ_B = 27.0
PUSH_HBM_WRITE_B = 24.7
# VALU instructions per push of push_kernel<float> (rocprofv3 SQ_INSTS_VALU of
# the all-zones launch x 64 lanes / 10M pushes;
# profiles/push_kernel_p2d-256x512-n1e7-3zones_sq_counters.txt) and MI355X's
# VALU issue rate: 256 CUs x 64 lanes x 2.4 GHz (one wave64 instruction per
# CU per cycle, non-packed).
PUSH_VALU_PER_PUSH = 484.0
# HBM bytes per particle of p3_push_kernel<float> on the 3-D workload (mean
# of the dispatches in profiles/p3_push_kernel_p3d-200x256x125-n5e7_traffic.txt,
# scaled by the same copy calibration)
P3_PUSH_HBM_READ_B = 36.4
P3_PUSH_HBM_WRITE_B = 41.6
# the same for push_kernel<float, true> on the 64-disc ensemble (mean of the
# dispatches in profiles/push_kernel_ens-64x128x256-n8.6e7_traffic.txt)
ENS_PUSH_HBM_READ_B = 29.7
ENS_PUSH_HBM_WRITE_B = 26.3
VALU_PEAK_LANE_OPS = 256 * 64 * 2.4e9
# algorithmic grid bytes of one C2 step besides the particles: the F_R and
# F_phi tables in FP32 (2 x 17 harmonics x 256 x 256 x 4 B, read once by the
# synthesis), the three zones' int64 mass sums read by the analysis
# (3 x 131072 x 8 B) and the FP32 (F_x, F_y) field written by the synthesis
# and read by the push (2 x 131072 x 8 B)
STEP_GRID_BYTES = 2 * 17 * 256 * 256 * 4 + 3 * 131072 * 8 + 2 * 131072 * 8
# adj_scatter_kernel<float>: algorithmic bytes per particle -- id (4), zone
# (1), (u, t) (8), the four FP64 adjoints read (32) and the two velocity
# adjoints written (16); the grid sums stay in LDS and L2.  Measured HBM
# bytes per particle: 41.1 read + 19.8 written
# (profiles/adj_scatter_kernel_p2d-128x256-n1e6-adjoint_traffic.txt).
ADJ_SCATTER_BYTES = 61.0
ADJ_SCATTER_HBM_B = 60.9


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def a2(x, y, lscale, r1=1.0, r2=3.0):
    """|A_2| = |sum e^{-2i phi}| / N over particles with r1 <= R <= r2."""
    x = x.astype(np.float64)
    y = y.astype(np.float64)
    R = np.hypot(x, y) / lscale
    sel = (R >= r1) & (R <= r2)
    ph = np.arctan2(y[sel], x[sel])
    c = np.exp(-2j * ph).sum() / max(sel.sum(), 1)
    return complex(c)


def run_gpu(w, precision, n, steps, warmup, sort_interval, graph=0, compensated=False):
    """Wall-clock rate over `steps` steps without events on the stream, then
    the same number of further steps with HIP events around every push
    launch (the push kernel's own time for the roofline)."""
    from galaxy_amd import configs
    t0 = time.time()
    g = configs.make_grid(w, precision=precision, n=n, compensated=compensated)
    g.set_sort_interval(sort_interval)
    g.set_graph(graph)
    log("[bench] %s set-up %.1f s (N=%d)" % (precision, time.time() - t0, n))
    if warmup:
        g.run(warmup)
    g.findf(True)                      # energy at the start (outside the timing)
    e0 = g.integrals()
    g.sync()
    g.set_profile(False)               # zero the push counter, no events
    t0 = time.perf_counter()
    g.run(steps)
    g.sync()
    t1 = time.perf_counter()
    _, _, pushes = g.get_profile()
    g.set_profile(True)
    g.run(steps)
    g.sync()
    push_ms, launches, pushes_p = g.get_profile()
    g.set_profile(False)
    g.findf(True)
    e1 = g.integrals()
    out = g.unload()
    res = dict(seconds=t1 - t0, pushes=pushes, push_ms=push_ms,
               launches=launches, pushes_prof=pushes_p,
               a2=a2(out["x"], out["y"], w.lscale()), noff=g.noff,
               e0=e0["energy"], e1=e1["energy"], lz0=e0["lz"], lz1=e1["lz"])
    log("[bench] %s: %d steps in %.4f s, %d pushes; profiled pass: push kernel "
        "%.3f ms over %d launches, %d pushes" % (precision, steps, t1 - t0, pushes,
                                                push_ms, launches, pushes_p))
    g.close()
    return res


def drift(r):
    """Relative energy change over the run (icheck's integrals, FP64 sums)."""
    return (r["e1"] - r["e0"]) / abs(r["e0"])


def accuracy_leg(w, n, steps, sort_interval):
    """The mixed and FP64 forms from the same initial particles, outside the
    timed region and independent of --steps: the relative force error of
    the mixed solve against the FP64 solve at the start (each form assigns
    its own mass; F_x, F_y and the potential, error over the grid's
    largest value), then `steps` steps of each (the workload's own length):
    the energy and Lz drift of each form, their difference, and the m = 2
    amplitude of both end states."""
    from galaxy_amd import configs
    from galaxy_amd.lib15 import P2DGrid
    t0 = time.time()
    rgrid = P2DGrid(precision="fp32", **configs.grid_kwargs(w, n)).rgrid
    parts = configs.particles(w, n, None, rgrid)
    fields, res = {}, {}
    for prec in ("fp32", "fp64"):
        g = P2DGrid(precision=prec, **configs.grid_kwargs(w, n))
        g.greenm()
        g.set_fixforce(*configs.fixforce(w, g.rgrid))
        g.set_fixpot(*configs.fixpot(w, g.rgrid))
        g.restrt(*parts)
        g.set_sort_interval(sort_interval)
        g.masset()
        g.findf(True)
        fields[prec] = g.get_field().astype(np.float64)
        e0 = g.integrals()
        g.run(steps)
        g.findf(True)
        e1 = g.integrals()
        out = g.unload()
        res[prec] = dict(drift=(e1["energy"] - e0["energy"]) / abs(e0["energy"]),
                         lz=(e1["lz"] - e0["lz"]) / abs(e0["lz"]),
                         a2=a2(out["x"], out["y"], w.lscale()))
        g.close()
    ferr = [float(np.abs(fields["fp32"][t] - fields["fp64"][t]).max()
                  / np.abs(fields["fp64"][t]).max()) for t in range(3)]
    m, d = res["fp32"], res["fp64"]
    log("[bench] accuracy leg: %d steps of both forms in %.1f s" % (steps, time.time() - t0))
    return dict(accuracy_steps=steps,
                force_err_vs_fp64=max(ferr), force_err_by_type=ferr,
                a2_mixed=abs(m["a2"]), a2_fp64=abs(d["a2"]), m2_err=abs(m["a2"] - d["a2"]),
                energy_drift_mixed=m["drift"], energy_drift_fp64=d["drift"],
                energy_drift_err=abs(m["drift"] - d["drift"]),
                lz_drift_mixed=m["lz"], lz_drift_fp64=d["lz"])


def cpu_baseline(w, n, steps):
    """The oracle (FP32 restatement, one core) on n particles of workload w."""
    from galaxy_amd import configs
    from oracle import oracle as orc
    kw = configs.grid_kwargs(w, n)
    m = orc.P2D(kw["nr"], kw["na"], kw["hash_"], kw["softl"], tsoft=1,
                nsect=1, nzones=kw["nzones"], nstep=kw["nstep"], zbr=kw["zbr"],
                pmass=kw["pmass"], offanal=True, prec="f", use_fx=True)
    m.fixforce = configs.fixforce(w, float(m.rgrid))
    m.load(*configs.particles(w, n, None, float(m.rgrid)))
    m.masset()
    pushes = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        m.findf()
        m.step()
        pushes += int(m.moved.sum())
    dt = time.perf_counter() - t0
    return dict(value=pushes / dt, unit="particle-steps/s", cores=1,
                kind="port",
                sample="oracle FP32, %s geometry, N=%d, %d steps (%d pushes, %.1f s)"
                       % (w.name, n, steps, pushes, dt))


def cpu_reference(w, n, steps):
    """The reference's own galaxy program (built from its sources into
    oracle/_ref, serial: one core) on n particles of workload w's disc at
    its grid, harmonics and zones; None when the build is absent.  Pushes
    per step from the zone populations of its first and last dumps."""
    import shutil
    import tempfile
    from oracle import refrun
    if not refrun.available():
        return None
    d = tempfile.mkdtemp(prefix="galaxy_ref_")
    try:
        refrun.write_run(d, w, n, steps)
        r = refrun.run_galaxy(d)
        per = refrun.pushes_per_step(d, 101, w.nstep)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    pushes = per * r["steps"]
    return dict(value=pushes / r["seconds"], unit="particle-steps/s", cores=1,
                kind="reference",
                sample="reference galaxy (flang build of SRC/, serial), %s geometry, N=%d, "
                       "%d steps, %.0f pushes per step, %.1f s by its own step lines "
                       "(set-up %.1f s outside; its Green's table re-read from disk every "
                       "solve is inside)" % (w.name, n, r["steps"], per, r["seconds"],
                                            r["setup_seconds"]))


def cpu_port_parallel(w, n, steps):
    """The oracle's particle loop on every host thread (OpenMP) at the full
    workload size, with the GPU's FP64 Green's tables (the oracle's own
    build of them takes minutes at this grid on one core)."""
    from galaxy_amd import configs
    from galaxy_amd.lib15 import P2DGrid
    from oracle import oracle as orc
    g = P2DGrid(precision="fp64", **configs.grid_kwargs(w, n))
    g.greenm()
    G = g.get_green()
    g.close()
    kw = configs.grid_kwargs(w, n)
    m = orc.P2D(kw["nr"], kw["na"], kw["hash_"], kw["softl"], tsoft=1, nsect=1,
                nzones=kw["nzones"], nstep=kw["nstep"], zbr=kw["zbr"], pmass=kw["pmass"],
                offanal=True, prec="f", use_fx=True, green_tables=G)
    m.fixforce = configs.fixforce(w, float(m.rgrid))
    m.load(*configs.particles(w, n, None, float(m.rgrid)))
    m.masset()
    pushes = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        m.findf()
        m.step(parallel=True)
        pushes += int(m.moved.sum())
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=pushes / dt, unit="particle-steps/s", cores=threads, kind="port",
                sample="oracle FP32 particle loop on %d OpenMP threads, %s, N=%d, %d steps "
                       "(%d pushes, %.1f s)" % (threads, w.name, n, steps, pushes, dt))


def run_gpu3(w, precision, n, steps, warmup, sort_interval, want_green=False):
    """The 3-D grid's step (findf -> p3fndf, step): wall-clock rate over
    `steps` steps, then a profiled pass of as many steps."""
    from galaxy_amd import configs
    t0 = time.time()
    g = configs.make_grid3(w, precision=precision, n=n)
    g.set_sort_interval(sort_interval)
    log("[bench] p3d %s set-up %.1f s (N=%d)" % (precision, time.time() - t0, n))
    g.set_fixpot(*configs.fixpot(w, math.hypot(g.rgrid, g.zm)))
    if warmup:
        g.run(warmup)
    g.findf(True)                      # energy at the start (outside the timing)
    e0 = g.integrals()
    g.sync()
    g.set_profile(False)
    t0 = time.perf_counter()
    g.run(steps)
    g.sync()
    t1 = time.perf_counter()
    _, _, pushes = g.get_profile()
    g.set_profile(True)
    g.run(steps)
    g.sync()
    push_ms, launches, pushes_p = g.get_profile()
    g.set_profile(False)
    g.findf(True)
    e1 = g.integrals()
    out = g.unload()
    res = dict(seconds=t1 - t0, pushes=pushes, push_ms=push_ms, launches=launches,
               pushes_prof=pushes_p, a2=a2(out["x"], out["y"], w.lscale()), noff=g.noff,
               e0=e0["energy"], e1=e1["energy"], lz0=e0["lz"], lz1=e1["lz"])
    if want_green:
        res["green"] = g.get_green()
    log("[bench] p3d %s: %d steps in %.4f s, %d pushes; profiled pass: push kernel "
        "%.3f ms over %d launches" % (precision, steps, t1 - t0, pushes, push_ms, launches))
    g.close()
    return res


def cpu_baseline3(w, n, steps, G):
    """The p3d oracle (FP32 particles, FP64 direct-sum solve, one core) on n
    particles of workload w, with the GPU's Green's tables."""
    from galaxy_amd import configs
    from oracle import oracle3
    kw = configs.grid_kwargs3(w, n)
    m = oracle3.P3D(kw["nr"], kw["na"], kw["ngz"], kw["dzg"], kw["hash_"], kw["softl"],
                    tsoft=1, nzones=kw["nzones"], nstep=kw["nstep"], zbr=kw["zbr"],
                    pmass=kw["pmass"], offanal=True, prec="f", green_tables=G)
    m.fixforce = configs.fixforce(w, math.hypot(float(m.rgrid), float(m.zm)))
    m.load(*configs.particles3(w, n, None, float(m.rgrid)))
    m.masset()
    pushes = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        m.findf()
        m.step()
        pushes += int(m.moved.sum())
    dt = time.perf_counter() - t0
    return dict(value=pushes / dt, unit="particle-steps/s", cores=1, kind="port",
                sample="oracle p3d (FP32 particles, FP64 direct-sum solve), %s geometry, "
                       "N=%d, %d steps (%d pushes, %.1f s)" % (w.name, n, steps, pushes, dt))


def main_p3d(args):
    """--workload C3: the 3-D polar grid, 200 x 256 x 125, N = 5e7."""
    from galaxy_amd import configs
    w = configs.C3
    n = args.n or w.n
    mixed = run_gpu3(w, "fp32", n, args.steps, args.warmup, args.sort_interval,
                     want_green=not args.no_cpu)
    fp64 = None
    if not args.no_fp64:
        fp64 = run_gpu3(w, "fp64", n, args.steps, args.warmup, args.sort_interval)
    value = mixed["pushes"] / mixed["seconds"]
    bytes_per_push = 48.0   # (x, y, z, vx) + (vy, vz) read and written, FP32
    achieved = mixed["pushes_prof"] * bytes_per_push / (mixed["push_ms"] * 1e-3) / 1e9
    per_launch = mixed["pushes_prof"] / max(mixed["launches"], 1)
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4),
                traffic=round((P3_PUSH_HBM_READ_B + P3_PUSH_HBM_WRITE_B) * per_launch),
                kernel="p3_push_kernel<float>", bytes_per_push=bytes_per_push,
                avg_launch_us=round(1e3 * mixed["push_ms"] / max(mixed["launches"], 1), 2))
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline3(w, args.cpu_n3, args.cpu_steps3, mixed.pop("green"))
        log("[bench] p3d cpu baseline: %.3g pushes/s (%s)" % (cpu["value"], cpu["sample"]))
    extra = dict(pushes_per_step=mixed["pushes"] / args.steps, sort_interval=args.sort_interval,
                 push_share=round(mixed["push_ms"] * 1e-3 / mixed["seconds"], 3), noff=mixed["noff"])
    if fp64:
        v64 = fp64["pushes"] / fp64["seconds"]
        extra.update(fp64_value=v64, mixed_over_fp64=value / v64,
                     fp64_ms_per_step=1e3 * fp64["seconds"] / args.steps,
                     a2_mixed=abs(mixed["a2"]), a2_fp64=abs(fp64["a2"]),
                     m2_err=abs(mixed["a2"] - fp64["a2"]),
                     energy_drift_mixed=drift(mixed), energy_drift_fp64=drift(fp64),
                     energy_drift_err=abs(drift(mixed) - drift(fp64)),
                     lz_drift_mixed=(mixed["lz1"] - mixed["lz0"]) / abs(mixed["lz0"]),
                     lz_drift_fp64=(fp64["lz1"] - fp64["lz0"]) / abs(fp64["lz0"]))
    line = {
        "metric": "particle-steps/s (3-D polar grid)", "value": value,
        "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * mixed["seconds"] / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded exponential disc, sech^2 vertical profile)",
        "config": {"workload": w.name, "grid": [w.nr, w.na, w.ngz], "n": n,
                   "harmonics": "m<=%d" % w.hash, "zones": list(w.nstep), "dt": w.dt,
                   "parallelism": "none"},
        "roofline": roof, "cpu_baseline": cpu, "extra": extra,
    }
    print(json.dumps(line), flush=True)


def run_gpu_ens(w, precision, scale, steps, warmup, sort_interval, graph, parts,
                contraction="exact"):
    """The ensemble workload: timed run (HIP graphs when graph is set), then
    a profiled pass of the same length with events around every push."""
    from galaxy_amd import configs
    t0 = time.time()
    g = configs.make_ensemble(w, precision=precision, scale=scale, parts=parts)
    g.set_sort_interval(sort_interval)
    g.set_graph(graph)
    if contraction == "mfma":
        g.set_contraction("mfma")
    log("[bench] ensemble %s set-up %.1f s (%d members, N=%d)"
        % (precision, time.time() - t0, g.nmem, g.n))
    if warmup:
        g.run(warmup)
    g.findf(True)
    e0 = [d["energy"] for d in g.integrals()]
    g.sync()
    g.set_profile(False)
    t0 = time.perf_counter()
    g.run(steps)
    g.sync()
    t1 = time.perf_counter()
    _, _, pushes = g.get_profile()
    g.set_profile(True)
    g.run(steps)
    g.sync()
    push_ms, launches, pushes_p = g.get_profile()
    g.set_profile(False)
    g.findf(True)
    e1 = [d["energy"] for d in g.integrals()]
    members = g.unload()
    a2s = [a2(m["x"], m["y"], w.lscale()) for m in members]
    res = dict(seconds=t1 - t0, pushes=pushes, push_ms=push_ms, launches=launches,
               pushes_prof=pushes_p, a2=a2s, e0=e0, e1=e1, n=g.n, nmem=g.nmem,
               contraction=g.get_contraction() if contraction == "mfma" else None)
    log("[bench] ensemble %s: %d steps in %.4f s, %d pushes; profiled pass: push kernel "
        "%.3f ms over %d launches" % (precision, steps, t1 - t0, pushes, push_ms, launches))
    g.close()
    return res


def main_ens(args):
    """--workload C4: 64 discs on 128 x 256 grids in one ensemble."""
    from galaxy_amd import configs, lib15
    w = configs.C4
    scale = 1.0 if not args.n else args.n / w.n
    counts, _ = w.members(scale)
    t0 = time.time()
    rgrid = lib15.P2DGrid(precision="fp32", **configs.grid_kwargs(w, 1)).rgrid
    parts = configs.ensemble_particles(w, counts, rgrid)
    log("[bench] ensemble particles drawn in %.1f s" % (time.time() - t0))
    mixed = run_gpu_ens(w, "fp32", scale, args.steps, args.warmup, args.sort_interval,
                        args.graph, parts, args.contraction)
    fp64 = None
    if not args.no_fp64:
        fp64 = run_gpu_ens(w, "fp64", scale, args.steps, args.warmup, args.sort_interval,
                           args.graph, parts)
    del parts
    value = mixed["pushes"] / mixed["seconds"]
    bytes_per_push = 32.0
    achieved = mixed["pushes_prof"] * bytes_per_push / (mixed["push_ms"] * 1e-3) / 1e9
    per_launch = mixed["pushes_prof"] / max(mixed["launches"], 1)
    traffic = None
    if ENS_PUSH_HBM_READ_B is not None:
        traffic = round((ENS_PUSH_HBM_READ_B + ENS_PUSH_HBM_WRITE_B) * per_launch)
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                kernel="push_kernel<float, true>", bytes_per_push=bytes_per_push,
                avg_launch_us=round(1e3 * mixed["push_ms"] / max(mixed["launches"], 1), 2))
    cpu = None
    if not args.no_cpu:
        i = int(np.argmax(counts))
        cpu = cpu_baseline(w.member(i, counts[i]), args.cpu_n, args.cpu_steps)
        cpu["sample"] = "member %d's disc: " % i + cpu["sample"]
        log("[bench] ensemble cpu baseline: %.3g pushes/s (%s)" % (cpu["value"], cpu["sample"]))
    extra = dict(members=mixed["nmem"], n_total=mixed["n"], graph=bool(args.graph),
                 sort_interval=args.sort_interval, contraction=args.contraction,
                 push_share=round(mixed["push_ms"] * 1e-3 / mixed["seconds"], 3))
    if mixed["contraction"] is not None:
        _, var, us = mixed["contraction"]
        extra.update(gemm_variant=int(var), gemm_variant_us=[round(float(u), 2) for u in us])
    if fp64:
        v64 = fp64["pushes"] / fp64["seconds"]
        d_m = [(b - a) / abs(a) for a, b in zip(mixed["e0"], mixed["e1"])]
        d_f = [(b - a) / abs(a) for a, b in zip(fp64["e0"], fp64["e1"])]
        extra.update(fp64_value=v64, mixed_over_fp64=value / v64,
                     fp64_ms_per_step=1e3 * fp64["seconds"] / args.steps,
                     m2_err_max=max(abs(a - b) for a, b in zip(mixed["a2"], fp64["a2"])),
                     a2_mixed_max=max(abs(a) for a in mixed["a2"]),
                     energy_drift_mixed_max=max(abs(v) for v in d_m),
                     energy_drift_err_max=max(abs(a - b) for a, b in zip(d_m, d_f)))
    line = {
        "metric": "particle-steps/s (ensemble of p2d discs)", "value": value,
        "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * mixed["seconds"] / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (64 seeded exponential discs, log-uniform counts, Q in [1, 2])",
        "config": {"workload": w.name, "members": w.nmem, "grid": [w.nr, w.na],
                   "n": mixed["n"], "harmonics": "m<=%d" % w.hash, "zones": list(w.nstep),
                   "dt": w.dt, "parallelism": "none"},
        "roofline": roof, "cpu_baseline": cpu, "extra": extra,
    }
    print(json.dumps(line), flush=True)


def main_adj(args):
    """--workload C5: the differentiated run (128 x 256 grid, N = 1e6, Q = 1.2,
    200 steps): |A_2|^2 at the end and its gradient in every initial
    velocity, Q and the disc mass by the adjoint kernels with checkpoints
    every 20 steps; the FP64 forward's gradient, the FP64 tangent-linear run
    and central differences along one random velocity direction as the
    checks."""
    from galaxy_amd import adjoint, configs
    w = configs.C5
    n = args.n or w.n
    t0 = time.time()
    m32 = adjoint.DiscModel(w, precision="fp32", n=n)
    log("[bench] adjoint model set-up %.1f s (N=%d)" % (time.time() - t0, n))
    if args.warmup:
        # one whole gradient: the model's grid, its adjoint slots and the
        # kernels' first launches are set up outside the timed call
        m32.gradient(w.Q, w.md, args.steps, every=20)
    t0 = time.perf_counter()
    g32 = m32.gradient(w.Q, w.md, args.steps, every=20)
    t32 = time.perf_counter() - t0
    log("[bench] adjoint fp32: %.3f s (%s)" % (t32, g32["seconds"]))
    # the same gradient again with HIP events around the adjoint particle
    # kernels (their own time for the roofline)
    gp = m32.gradient(w.Q, w.md, args.steps, every=20, profile=True)
    sc_ms, ga_ms, nprof = gp["kernel_ms"]
    log("[bench] adjoint profiled pass: scatter %.3f ms, gather %.3f ms over %d steps"
        % (sc_ms, ga_ms, nprof))
    bpp = ADJ_SCATTER_BYTES
    achieved = n * nprof * bpp / (sc_ms * 1e-3) / 1e9
    traffic = None
    if ADJ_SCATTER_HBM_B is not None:
        traffic = round(ADJ_SCATTER_HBM_B * n)
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                kernel="adj_scatter_kernel<float>", bytes_per_particle=bpp,
                avg_launch_us=round(1e3 * sc_ms / max(nprof, 1), 2),
                gather_avg_launch_us=round(1e3 * ga_ms / max(nprof, 1), 2))
    extra = dict(loss=g32["loss"], dQ=g32["dQ"], dmd=g32["dmd"], every=20,
                 seconds_fp32={k: round(v, 4) for k, v in g32["seconds"].items()})
    if not args.no_fp64:
        m64 = adjoint.DiscModel(w, precision="fp64", n=n)
        t0 = time.perf_counter()
        g64 = m64.gradient(w.Q, w.md, args.steps, every=20, with_sel=True)
        t64 = time.perf_counter() - t0
        v32 = np.concatenate([g32["dvx"], g32["dvy"]])
        v64 = np.concatenate([g64["dvx"], g64["dvy"]])
        rng = np.random.default_rng(17)
        dx, dy = rng.standard_normal(n), rng.standard_normal(n)
        vx, vy = m64.velocities(w.Q, w.md)
        vs = float(np.sqrt(np.mean(vx * vx + vy * vy)))
        ad = float(np.dot(g64["dvx"], dx) + np.dot(g64["dvy"], dy))
        # the FP64 tangent-linear run along the same direction: an
        # independent forward-mode reference for the adjoint's number
        d0 = np.zeros((4, n))
        d0[2], d0[3] = dx, dy
        t0 = time.perf_counter()
        tl, _, _ = m64.tangent(w.Q, w.md, args.steps, d0)
        t_tl = time.perf_counter() - t0
        fds = []
        # h = 1e-4 |v| (the config's step) crosses CIC cell edges: the map is
        # piecewise smooth in the positions, so the difference there measures
        # the kinks as much as the derivative
        for hr in (1e-4, 1e-6, 1e-7):
            h = hr * vs
            # the annulus membership of the unperturbed end state held fixed,
            # as the adjoint holds it (at this N single particles crossing
            # R = 1 or 3 would otherwise dominate the difference)
            lp = m64.loss_v(vx + h * dx, vy + h * dy, w.md, args.steps, sel=g64["sel"])
            lm = m64.loss_v(vx - h * dx, vy - h * dy, w.md, args.steps, sel=g64["sel"])
            fd = (lp - lm) / (2 * h)
            fds.append(dict(h_rel=hr, central_difference=fd, rel_err=abs(ad - fd) / abs(fd),
                            rel_err_tangent=abs(tl - fd) / abs(fd)))
        extra.update(fp64_seconds=round(t64, 3), loss_fp64=g64["loss"],
                     grad_v_rel_err_fp32_vs_fp64=float(np.linalg.norm(v32 - v64)
                                                       / np.linalg.norm(v64)),
                     dQ_fp64=g64["dQ"], dmd_fp64=g64["dmd"],
                     fd_direction=dict(adjoint=ad, tangent_linear=tl,
                                       tangent_vs_adjoint=abs(tl - ad) / abs(ad),
                                       tangent_seconds=round(t_tl, 3), checks=fds))
    value = n * args.steps / t32
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(w, args.cpu_n // 2, args.cpu_steps)
        cpu["sample"] = "forward only (no CPU adjoint exists): " + cpu["sample"]
    line = {
        "metric": "particle-steps/s of the differentiated run (forward + adjoint)",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * t32 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded exponential disc, Q = 1.2)",
        "config": {"workload": w.name, "grid": [w.nr, w.na], "n": n,
                   "harmonics": "m<=%d" % w.hash, "dt": w.dt, "loss": "|A2|^2, R in [1, 3]",
                   "checkpoint_every": 20, "parallelism": "none"},
        "roofline": roof, "cpu_baseline": cpu, "extra": extra,
    }
    print(json.dumps(line), flush=True)


def copy_calibration(nbytes):
    """A plain elementwise copy of known size (FP32 x 1.0 into a second
    buffer: nbytes read, nbytes written), run three times; its FETCH_SIZE and
    WRITE_SIZE per dispatch scale the same counters of the push kernel."""
    import torch
    n = nbytes // 4
    src = torch.rand(n, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": 4 * n, "dispatches": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=None, help="particles (default 10M)")
    ap.add_argument("--sort-interval", type=int, default=None,
                    help="steps between re-sorts (default 10 for C2, 5 for C3, 15 for C4)")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 leg")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the accuracy leg (C2: both forms over the workload's 500 steps)")
    ap.add_argument("--cpu-n", type=int, default=2_000_000)
    ap.add_argument("--cpu-ref-n", type=int, default=1_000_000,
                    help="particles of the reference's own CPU run (C2)")
    ap.add_argument("--cpu-ref-steps", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain copy kernel of BYTES (counter calibration)")
    ap.add_argument("--workload", default="C2", choices=("C2", "C3", "C4", "C5"),
                    help="C2: 2-D grid, the headline (default); C3: 3-D grid; "
                         "C4: ensemble of 64 2-D discs")
    ap.add_argument("--contraction", default=None, choices=("exact", "mfma"),
                    help="ensemble radial contraction: FP32 dot products in the synthesis "
                         "(exact) or split-BF16 GEMM on MFMA (default for C4)")
    ap.add_argument("--compensated", action="store_true",
                    help="mixed form with compensated position updates (C2)")
    ap.add_argument("--graph", type=int, default=None, choices=(0, 1),
                    help="run the timed steps as HIP graphs (default: 1 for C4, 0 otherwise)")
    ap.add_argument("--cpu-n3", type=int, default=4_000_000)
    ap.add_argument("--cpu-steps3", type=int, default=2)
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (the step's "
              "multi-GPU form is a design, not measured); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "C5":
        if args.steps == 500:
            args.steps = 200
        main_adj(args)
        return
    if args.workload == "C4":
        if args.steps == 500:
            args.steps = 300
        if args.sort_interval is None:
            args.sort_interval = 20
        if args.graph is None:
            args.graph = 1
        if args.contraction is None:
            args.contraction = "mfma"
        main_ens(args)
        return
    if args.workload == "C3":
        if args.steps == 500:
            args.steps = 200
        if args.sort_interval is None:
            args.sort_interval = 8
        main_p3d(args)
        return
    if args.sort_interval is None:
        args.sort_interval = 20
    from galaxy_amd import configs
    w = configs.C2
    n = args.n or w.n
    graph = args.graph or 0
    mixed = run_gpu(w, "fp32", n, args.steps, args.warmup, args.sort_interval, graph,
                    args.compensated)
    fp64 = None
    if not args.no_fp64:
        fp64 = run_gpu(w, "fp64", n, args.steps, args.warmup, args.sort_interval, graph)
    comp = None
    if not args.compensated and not args.no_fp64:
        # the same mixed form with compensated position updates, beside it
        comp = run_gpu(w, "fp32", n, args.steps, args.warmup, args.sort_interval, graph, True)
    value = mixed["pushes"] / mixed["seconds"]
    ms_per_step = 1e3 * mixed["seconds"] / args.steps
    bytes_per_push = 32.0
    achieved = mixed["pushes_prof"] * bytes_per_push / (mixed["push_ms"] * 1e-3) / 1e9
    per_launch = mixed["pushes_prof"] / max(mixed["launches"], 1)
    traffic = (PUSH_HBM_READ_B + PUSH_HBM_WRITE_B) * per_launch
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS,
                unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4),
                traffic=round(traffic),
                kernel="push_kernel<float>",
                bytes_per_push=bytes_per_push,
                avg_launch_us=round(1e3 * mixed["push_ms"] / max(mixed["launches"], 1), 2))
    cpu = None
    cpu_par = None
    if not args.no_cpu:
        cpu = cpu_reference(w, args.cpu_ref_n, args.cpu_ref_steps)
        if cpu is None:
            cpu = cpu_baseline(w, args.cpu_n, args.cpu_steps)
        log("[bench] cpu baseline: %.3g pushes/s (%s)" % (cpu["value"], cpu["sample"]))
        cpu_par = cpu_port_parallel(w, n, 4)
        log("[bench] cpu port, all threads: %.3g pushes/s (%s)"
            % (cpu_par["value"], cpu_par["sample"]))
    in_kernel = mixed["pushes_prof"] / (mixed["push_ms"] * 1e-3)
    # the step-level fraction: (pushes x 32 B + the two field types' FP32
    # tables + zone mass sums + field written and read) per step over the
    # wall-clock step time
    step_bytes = (mixed["pushes"] / args.steps * bytes_per_push + STEP_GRID_BYTES)
    roof["step_frac"] = round(step_bytes / (ms_per_step * 1e-3) / (HBM_PEAK_GBS * 1e9), 4)
    roof["step_bytes"] = round(step_bytes)
    extra = dict(
        n_steps_rate=n * args.steps / mixed["seconds"],
        push_valu_frac=round(in_kernel * PUSH_VALU_PER_PUSH / VALU_PEAK_LANE_OPS, 3),
        pushes_per_step=mixed["pushes"] / args.steps,
        sort_interval=args.sort_interval,
        graph=bool(graph),
        compensated=bool(args.compensated),
    )
    if fp64:
        v64 = fp64["pushes"] / fp64["seconds"]
        extra.update(fp64_value=v64, mixed_over_fp64=value / v64,
                     fp64_ms_per_step=1e3 * fp64["seconds"] / args.steps)
    if comp:
        extra.update(compensated_value=comp["pushes"] / comp["seconds"],
                     compensated_ms_per_step=1e3 * comp["seconds"] / args.steps,
                     compensated_push_us=round(1e3 * comp["push_ms"] / max(comp["launches"], 1),
                                               2))
    if cpu_par:
        extra["cpu_port_parallel"] = cpu_par
    if not args.no_accuracy:
        extra.update(accuracy_leg(w, n, w.steps, args.sort_interval))
    line = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded exponential disc)",
        "config": {"workload": w.name, "grid": [w.nr, w.na], "n": n,
                   "harmonics": "m<=%d" % w.hash, "zones": list(w.nstep),
                   "dt": [w.dt * s for s in w.nstep], "parallelism": "none"},
        "roofline": roof, "cpu_baseline": cpu, "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
