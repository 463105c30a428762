"""The oracle against the reference's own programs, built from its sources
with this image's flang into oracle/_ref (oracle/Makefile.ref).

  - The build is faithful: 500 steps of galaxy from the shipped
    EXAMPLES/p2d/run100.dmp0 with the shipped run100.dat reach the shipped
    run100.dmp10 (written by the reference author's own build) to 1e-6,
    every particle on the same zone list.
  - The log-spiral coefficients the reference's analysis writes (its .res
    LGSP records, lgspia.f / phyprp.f) at steps 0 and 500 equal the
    oracle's on the dumps of those steps to its FP32 summation error.
  - The reference's FP32 Green's tables (the .grt file its p2grnm writes,
    records in p2gwrt's order: F_R for m = 0..ng-1, F_phi for m = 1..,
    the potential) equal the oracle's FP64 tables to its FP32 rounding.
  - The oracle's step against the reference's on the benchmark discs: the
    exponential disc with its fixed flat rotation curve (a rigid halo of
    the reference's FE type, whose force is the one the workloads tabulate)
    at the 128 x 128 grid over 20 steps, and at the 256 x 512 grid with its
    three time-step zones over 40 steps (the zone schedule, rezgrp, chstep
    and zoncmb as the reference runs them).  Both start from the
    reference's own pcs2dmp output; positions agree to FP32 rounding
    (the reference's FP32 FFT against the oracle's FP64 solve), zones
    exactly.  The pcs2dmp start itself (load, zones, masset, findf,
    tmcent) is checked the same way.

CPU only; skipped where the reference tree or its build is absent.
"""
import os
import shutil

import numpy as np
import pytest

from galaxy_amd import configs, dmp, pcs

EX = "/root/reference/EXAMPLES/p2d"

try:
    from oracle import refrun
    HAVE_REF = refrun.available() and os.path.isdir(EX)
except Exception:  # pragma: no cover
    HAVE_REF = False

pytestmark = pytest.mark.skipif(not HAVE_REF, reason="reference build (oracle/_ref) absent")


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.fixture(scope="module")
def run100(tmp_path_factory):
    """galaxy, 500 steps from the shipped run100.dmp0."""
    d = str(tmp_path_factory.mktemp("run100"))
    refrun.galaxy_from_dump(d, 100, os.path.join(EX, "run100.dmp0"),
                            os.path.join(EX, "run100.dat"), 500)
    return d


def test_build_reproduces_shipped_dump(run100):
    a = dmp.read_dmp(os.path.join(run100, "run100.dmp10"))
    b = dmp.read_dmp(os.path.join(EX, "run100.dmp10"))
    assert a["istep"] == b["istep"] == 500
    for k in ("x", "y", "vx", "vy"):
        assert _rel(a[k], b[k]) <= 1e-6, k
    assert np.array_equal(a["iz"], b["iz"])


def test_green_tables_match_oracle(run100, orc):
    recs = dmp._records(open(os.path.join(run100, "run100.grt"), "rb").read())
    nr, na, ng, nsect = 93, 128, 3, 2
    mats = [np.frombuffer(r, "<f4").reshape(nr, nr) for r in recs[1:]]
    G, _ = orc.green(nr, na, ng, nsect, orc.alpha(na), 3.67879, 1)
    order = [(0, m) for m in range(ng)] + [(1, m) for m in range(1, ng)] + \
            [(2, m) for m in range(ng)]
    assert len(mats) == len(order)
    for (t, m), A in zip(order, mats):
        B = G[t, m]
        if not B.any():
            assert not A.any(), (t, m)
            continue
        assert _rel(A, B) <= 1e-5, (t, m, _rel(A, B))


def _reference_run(tmp_path, w, n, steps):
    d = str(tmp_path / w.name)
    coords = refrun.write_run(d, w, n, steps)
    refrun.run_galaxy(d)
    return d, coords


def _oracle_from_dump(orc, w, d0, n):
    """The oracle at the reference's start state (FP32, the run's grid)."""
    kw = configs.grid_kwargs(w, n)
    ls, ts = np.float32(w.lscale()), np.float32(w.dt)
    pm = pcs.internal_mass(np.float32(w.md / n), float(ls), float(ts))
    zbr = tuple(float(np.float32(z)) for z in kw["zbr"])
    m = orc.P2D(kw["nr"], kw["na"], kw["hash_"], float(np.float32(kw["softl"])),
                nzones=kw["nzones"], nstep=kw["nstep"], zbr=zbr, pmass=pm, prec="f",
                use_fx=True)
    m.fixforce = configs.fixforce(w, float(m.rgrid))
    m.load(d0["x"], d0["y"], d0["vx"], d0["vy"], d0["iz"])
    return m, pm


@pytest.mark.parametrize("name,n,steps,tol", [("C1", 20000, 20, 2e-6),
                                              ("C2", 20000, 40, 2e-6)])
def test_oracle_matches_reference_steps(tmp_path, orc, name, n, steps, tol):
    w = configs.WORKLOADS[name]
    d, coords = _reference_run(tmp_path, w, n, steps)
    d0 = dmp.read_dmp(os.path.join(d, "run101.dmp0"))
    d1 = dmp.read_dmp(os.path.join(d, "run101.dmp1"))
    assert d1["istep"] == steps
    # the pcs2dmp start: FP32 unit conversion, zones, step factors, masset,
    # a full findf, tmcent
    m, pm = _oracle_from_dump(orc, w, d0, n)
    x, y, vx, vy = pcs.to_internal(coords, w.lscale(), w.dt)
    iz = pcs.zones(x, y, tuple(np.float32(z) for z in m.zbr), float(m.rgrid))
    assert np.array_equal(iz, d0["iz"])
    vx, vy = pcs.adjvls(vx, vy, iz, w.nstep)
    s, _ = _oracle_from_dump(orc, w, d0, n)
    s.load(x, y, vx, vy, iz)
    s.masset()
    s.findf()
    s.tmcent(True)
    assert np.array_equal(s.x, d0["x"]) and np.array_equal(s.y, d0["y"])
    for k in ("vx", "vy"):
        assert _rel(getattr(s, k), d0[k]) <= 1e-6, ("start", k)
    # the steps
    m.masset()
    for _ in range(steps):
        m.findf()
        m.step()
    errs = {k: _rel(getattr(m, k), d1[k]) for k in ("x", "y", "vx", "vy")}
    print(name, errs, "off", m.noff, d1["noff"])
    for k in ("x", "y"):
        assert errs[k] <= tol, (k, errs[k])
    for k in ("vx", "vy"):
        assert errs[k] <= 5 * tol, (k, errs[k])
    assert np.array_equal(m.iz, d1["iz"])
    shutil.rmtree(d, ignore_errors=True)


def test_lgspia_matches_reference(run100, orc):
    """LGSP records of the run's .res (ipp, 'LGSP', istep, np, nm; then the
    cos and sin parts, phyprp.f:182-197) at steps 0 and 500 against the
    oracle on the particles of the shipped dmp0 and of the run's dmp10
    (run100: two sectors, lscale 10, offgrid t).  The values are ~1e-6 (a
    quiet start); the reference's FP32 sums of 15000 unit vectors carry
    ~1e-8, the tolerance is 5e-8 absolute."""
    import struct
    recs = dmp._records(open(os.path.join(run100, "run100.res"), "rb").read())
    lg = {}
    for i, r in enumerate(recs):
        if len(r) == 20 and r[4:8] == b"LGSP":
            _, istep, np_, nm = struct.unpack("<i4xiii", r)
            lg[istep] = (np_, nm, np.frombuffer(recs[i + 1], "<f4").astype(np.float64))
    for istep, path in ((0, os.path.join(EX, "run100.dmp0")),
                        (500, os.path.join(run100, "run100.dmp10"))):
        np_, nm, v = lg[istep]
        d = dmp.read_dmp(path)
        assert d["istep"] == istep
        c, s = orc.lgspia(d["x"], d["y"], d["iz"], d["nz"], True, 2, np_, nm, 10.0)
        k = np_ * nm
        assert np.abs(v).max() > 1e-6
        assert np.abs(c - v[:k]).max() <= 5e-8, istep
        assert np.abs(s - v[k:]).max() <= 5e-8, istep
