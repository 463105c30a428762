"""The five benchmark workloads on the GPU, each at its own geometry.

Each workload runs the HIP path through the C ABI at the grid, harmonic
filter, zones and disc it is defined with (galaxy_amd/configs.py), against
the oracle at particle counts the oracle finishes in seconds, and at the
full particle count through properties that do not need the oracle:

  p2d-128x128-n1e5 (FP64): 100 steps of the FP64 GPU run against the FP64
    oracle's own solve and step: zones, cells and per-cell counts exact,
    positions within 1e-9 of the grid radius.
  p2d-256x512-n1e7-3zones: at N = 2e5, a full zone cycle (mzone 3, 1, 2, 1)
    with the oracle's field injected is bit-exact in FP32; the FP32 solve is
    within 1e-5 of the FP64 oracle's (relative force error).  At N = 1e7,
    40 steps of the mixed and FP64 runs: ncheck's particle count at every
    full step (SRC/lib15/ncheck.f:30-85), the zone ring ranges against the
    particles' cells, bounded energy drift in both forms and their
    difference, and the m = 2 amplitude of the two forms within 1e-4.
  p3d-200x256x125-n5e7: at N = 2e5, two steps with the oracle's field
    injected, bit-exact in FP32, and the FP32 solve within 1e-5 of the FP64
    oracle's; at N = 5e7, 10 steps: particle count, ranges, energy drift.
  ens-64x(128x256): the 64 members at 1 % of their counts: one member's
    solve and step against the oracle directly (FP32 bit-exact step with the
    oracle's field), every member bit-identical to its lone disc.
  p2d-128x256-n1e6-adjoint: the 200-step gradient at N = 1e6 against
    central differences of the FP64 forward run, and the transposed solve
    at this geometry.
"""
import numpy as np
import pytest

from galaxy_amd import configs
from tests._model import gpu_model, oracle_model, rel_err
from tests.test_gpu_parity import _chain, _close, _same, _say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def green_c2(orc):
    """The FP64 oracle's own Green's tables at 256 x 512, m <= 16 (built
    once: about ten seconds on one core)."""
    w = configs.C2
    kw = configs.grid_kwargs(w, 1)
    G, _ = orc.green(w.nr, w.na, w.hash + 1, 1, orc.alpha(w.na), kw["softl"], 1)
    return G


# ---- p2d-128x128-n1e5, FP64 -------------------------------------------------

def test_c1_fp64_100_steps(orc):
    w = configs.C1
    _say("C1", w.name, "fp64, N = %d, %d steps" % (w.n, w.steps))
    m, parts = oracle_model(orc, w, prec="d")
    g = gpu_model(w, precision="fp64", parts=parts)
    g.set_sort_interval(10)
    m.masset()
    g.masset()
    for _ in range(w.steps):
        m.findf()
        m.step()
    g.run(w.steps)
    out = g.unload()
    _same(out["iz"], m.iz, "zones")
    scale = float(g.rgrid)
    dx = max(np.abs(out[k] - getattr(m, k)).max() for k in ("x", "y")) / scale
    _say("C1 max position difference / rgrid = %.3e" % dx)
    assert dx <= 1e-9, dx
    on = out["iz"] <= w.nzones
    c_g, _, _ = orc.weight(out["x"][on], out["y"][on], w.nr, w.na, "d")
    c_o, _, _ = orc.weight(m.x[on], m.y[on], w.nr, w.na, "d")
    _same(c_g, c_o, "cells")
    _same(np.bincount(c_g, minlength=g.mesh + g.na), np.bincount(c_o, minlength=g.mesh + g.na),
          "per-cell counts")
    assert g.noff == m.noff
    mfx, _, jr, kr, qb = g.get_mass()
    _same(jr, m.jrad, "jrad")
    _same(kr, m.krad, "krad")
    _close(mfx * 2.0 ** -qb, m.mfx * 2.0 ** -qb, 1e-7, "mass sums")
    g.close()


# ---- p2d-256x512-n1e7-3zones -------------------------------------------------

def test_c2_zone_cycle_bit_exact(orc, green_c2):
    """N = 2e5 on the workload's grid, harmonics and zones: four steps
    (mzone 3, 1, 2, 1) plus one, the oracle's field injected, FP32
    bit-exact (positions, velocities, zones, mass sums, ranges)."""
    w = configs.C2
    _say("C2 zone cycle fp32, N = 2e5")
    m, parts = oracle_model(orc, w, prec="f", n=200_000, G=green_c2)
    g = gpu_model(w, precision="fp32", n=200_000, parts=parts)
    _chain(m, g, 5, True)
    g.close()


def test_c2_force_error_vs_fp64_oracle(orc, green_c2):
    """The mixed form's field against the FP64 oracle's on the same initial
    particles (each side assigns its own mass): relative force error
    <= 1e-5 for F_x, F_y and the potential, full and partial ranges."""
    w = configs.C2
    _say("C2 force error, N = 2e5")
    m, parts = oracle_model(orc, w, prec="d", n=200_000, G=green_c2)
    g = gpu_model(w, precision="fp32", n=200_000, parts=parts)
    m.masset()
    g.masset()
    errs = []
    for lpot in (True, False):
        m.findf(lpot)
        g.findf(lpot)
        assert g.ranges() == m.ranges
        f = g.get_field().astype(np.float64)
        for t in range(3 if lpot else 2):
            errs.append(rel_err(f[t], m.fld[t]))
    _say("C2 relative force error: %s" % ", ".join("%.2e" % e for e in errs))
    assert max(errs) <= 1e-5, errs
    g.close()


def _a2(x, y, lscale, r1=1.0, r2=3.0):
    """A_2 = sum e^{-2 i phi} / count over the particles with r1 <= R <= r2
    (FP64 on the host)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    R = np.hypot(x, y) / lscale
    s = (R >= r1) & (R <= r2)
    return complex(np.exp(-2j * np.arctan2(y[s], x[s])).sum() / max(s.sum(), 1))


def _ring_ranges(orc, w, out, nz, op):
    """jrad, krad from the particles' cells (massgn.f:80-89)."""
    jr = np.full(nz, w.nr, np.int32)
    kr = np.zeros(nz, np.int32)
    for z in range(1, nz + 1):
        s = out["iz"] == z
        if not s.any():
            continue
        c, _, _ = orc.weight(out["x"][s], out["y"][s], w.nr, w.na, op)
        ju = c // w.na
        jr[z - 1] = ju.min() + 1
        kr[z - 1] = ju.max() + 2
    return jr, kr


def test_c2_full_size_properties(orc):
    """N = 1e7, 40 steps in both forms: ncheck's count at each full step,
    ring ranges, energy drift, and the m = 2 amplitude of mixed vs FP64."""
    w = configs.C2
    _say("C2 full size, N = %d" % w.n)
    res = {}
    for prec in ("fp32", "fp64"):
        g = configs.make_grid(w, precision=prec)
        g.set_sort_interval(10)
        g.findf(True)
        e0 = g.integrals()
        for k in range(10):
            g.run(4)          # ends on a full step (nstep of the last zone = 4)
            mfx, _, jr, kr, qb = g.get_mass()
            tot = float((mfx.astype(np.float64) * 2.0 ** -qb).sum())
            assert round(tot) == w.n - g.noff, (prec, k, tot, g.noff)
        out = g.unload()
        ej, ek = _ring_ranges(orc, w, out, w.nzones, "f" if prec == "fp32" else "d")
        _same(jr, ej, prec + " jrad")
        _same(kr, ek, prec + " krad")
        g.findf(True)
        e1 = g.integrals()
        drift = (e1["energy"] - e0["energy"]) / abs(e0["energy"])
        res[prec] = (drift, _a2(out["x"], out["y"], w.lscale()))
        _say("C2 %s: energy drift %.3e over 40 steps, |A2| %.3e, %d off the grid"
             % (prec, drift, abs(res[prec][1]), g.noff))
        assert abs(drift) < 1e-4
        g.close()
    assert abs(res["fp32"][0] - res["fp64"][0]) < 1e-7
    assert abs(res["fp32"][1] - res["fp64"][1]) < 1e-4


# ---- p3d-200x256x125-n5e7 ----------------------------------------------------

@pytest.fixture(scope="module")
def c3_fp64_tables():
    """The FP64 Green's tables of the 3-D workload, built on the GPU (they
    are held to the oracle's to 1e-12 on the small grid)."""
    from galaxy_amd.lib15 import P3DGrid
    g = P3DGrid(precision="fp64", **configs.grid_kwargs3(configs.C3, 1))
    g.greenm()
    G = g.get_green()
    g.close()
    return G


def _models3(prec, op, n, G):
    import math
    from galaxy_amd.lib15 import P3DGrid
    from oracle import oracle3
    w = configs.C3
    kw = configs.grid_kwargs3(w, n)
    g = P3DGrid(precision=prec, **kw)
    g.set_green(G)
    m = oracle3.P3D(kw["nr"], kw["na"], kw["ngz"], kw["dzg"], kw["hash_"], kw["softl"],
                    tsoft=1, nzones=kw["nzones"], nstep=kw["nstep"], zbr=kw["zbr"],
                    pmass=kw["pmass"], offanal=True, prec=op, green_tables=G)
    ff = configs.fixforce(w, math.hypot(float(g.rgrid), float(g.zm)))
    g.set_fixforce(*ff)
    m.fixforce = ff
    parts = configs.particles3(w, n, None, g.rgrid)
    g.restrt(*parts)
    m.load(*parts)
    return m, g


def test_c3_two_steps_bit_exact(c3_fp64_tables):
    """N = 2e5 on the 200 x 256 x 125 grid: masset, the FP32 solve within
    1e-5 of the FP64 oracle's, and two steps with the oracle's field
    injected, bit-exact in FP32."""
    from tests.test_gpu_p3d import chain3
    _say("C3 two steps fp32, N = 2e5")
    m, g = _models3("fp32", "f", 200_000, c3_fp64_tables)
    chain3(m, g, 2, True)
    g.close()


def test_c3_force_error_vs_fp64_oracle(c3_fp64_tables):
    _say("C3 force error, N = 2e5")
    m, _ = _models3("fp64", "d", 200_000, c3_fp64_tables)
    _, g = _models3("fp32", "f", 200_000, c3_fp64_tables)
    m.masset()
    g.masset()
    m.findf(True)
    g.findf(True)
    assert g.ranges() == m.ranges
    f = g.get_field().astype(np.float64)
    errs = [np.abs(f[t] - m.fld[t]).max() / np.abs(m.fld[t]).max() for t in range(4)]
    _say("C3 relative force error: %s" % ", ".join("%.2e" % e for e in errs))
    assert max(errs) <= 1e-5, errs
    g.close()


def test_c3_full_size_properties():
    """N = 5e7, 10 steps (FP32): every particle accounted for (on-grid mass
    = N - noff), energy drift bounded, no non-finite coordinate."""
    w = configs.C3
    _say("C3 full size, N = %d" % w.n)
    g = configs.make_grid3(w, precision="fp32")
    import math
    g.set_fixpot(*configs.fixpot(w, math.hypot(g.rgrid, g.zm)))
    g.set_sort_interval(5)
    g.findf(True)
    e0 = g.integrals()
    g.run(10)
    mfx, _, jr, kr, qb = g.get_mass()
    tot = float((mfx.astype(np.float64) * 2.0 ** -qb).sum())
    assert round(tot) == w.n - g.noff, (tot, g.noff)
    assert jr[0] >= 1 and kr[0] <= w.nr and jr[0] < kr[0], (jr, kr)
    g.findf(True)
    e1 = g.integrals()
    drift = (e1["energy"] - e0["energy"]) / abs(e0["energy"])
    _say("C3 energy drift %.3e over 10 steps, %d off the grid" % (drift, g.noff))
    assert abs(drift) < 1e-4
    out = g.unload()
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.isfinite(out[k]).all(), k
    g.close()


# ---- ens-64x(128x256) --------------------------------------------------------

def test_c4_members(orc):
    """The 64-disc ensemble at 1 % of its member counts (Q_i, seeds and
    grid as defined): member 0 against the oracle (solve within 1e-5 of
    the FP64 oracle; three steps with the oracle's field injected into it,
    bit-exact), and after 12 more steps every member bit-identical to its
    lone disc."""
    from galaxy_amd.lib15 import P2DGrid
    w = configs.C4
    scale = 0.01
    counts, Q = w.members(scale)
    _say("C4 ensemble, %d members, N = %d" % (w.nmem, counts.sum()))
    kw = configs.grid_kwargs(w, 1)
    kw.pop("pmass")
    rgrid = P2DGrid(precision="fp32", pmass=1.0, **kw).rgrid
    parts = configs.ensemble_particles(w, counts, rgrid, workers=1)
    E = configs.make_ensemble(w, precision="fp32", scale=scale, parts=parts)
    E.set_sort_interval(5)
    i = 0
    mw = w.member(i, counts[i])
    mo = orc.P2D(kw["nr"], kw["na"], kw["hash_"], kw["softl"], tsoft=1, nsect=1,
                 nzones=kw["nzones"], nstep=kw["nstep"], zbr=kw["zbr"],
                 pmass=configs.grid_kwargs(mw, counts[i])["pmass"], offanal=True,
                 prec="f", use_fx=True)
    mo.fixforce = configs.fixforce(mw, float(mo.rgrid))
    md = orc.P2D(kw["nr"], kw["na"], kw["hash_"], kw["softl"], tsoft=1, nsect=1,
                 nzones=kw["nzones"], nstep=kw["nstep"], zbr=kw["zbr"],
                 pmass=configs.grid_kwargs(mw, counts[i])["pmass"], offanal=True,
                 prec="d", use_fx=True, green_tables=mo.G)
    md.fixforce = mo.fixforce
    mo.load(*parts[i])
    md.load(*parts[i])
    mo.masset()
    md.masset()
    _same(E.get_mass(i)[0], mo.mfx, "member 0 mass sums")
    E.findf()
    md.findf()
    err = max(rel_err(E.get_field(i)[t], md.fld[t]) for t in range(2))
    _say("C4 member 0 relative force error vs the FP64 oracle %.2e" % err)
    assert err <= 1e-5, err
    for k in range(3):
        if k:
            E.findf()
        mo.findf()
        assert E.ranges(i) == mo.ranges
        err = max(rel_err(E.get_field(i)[t], mo.fld[t]) for t in range(2))
        assert err <= 1e-5, (k, err)
        fld = mo.fld.astype(np.float32)
        E.set_field(i, fld)
        mo.step(field=fld[:2])
        E.step()
        o = E.unload()[i]
        for key in ("x", "y", "vx", "vy", "iz"):
            _same(o[key], getattr(mo, key), "member 0 %s step %d" % (key, k))
        _same(E.get_mass(i)[0], mo.mfx, "member 0 mass sums step %d" % k)
    # every member against its lone disc, from the same (now advanced) state
    members = E.unload()
    G = E.get_green()
    lone = []
    for j in range(w.nmem):
        nj = max(int(counts[j]), 1)
        S = P2DGrid(precision="fp32", pmass=configs.grid_kwargs(w.member(j, nj), nj)["pmass"],
                    **kw)
        S.set_green(G)
        S.set_fixforce(*configs.fixforce(w, S.rgrid))
        mj = members[j]
        S.restrt(mj["x"], mj["y"], mj["vx"], mj["vy"], mj["iz"])
        S.set_sort_interval(5)
        lone.append(S)
    for gg in [E] + lone:
        gg.masset()
        gg.run(12)
    members = E.unload()
    for j, S in enumerate(lone):
        o = S.unload()
        for key in ("x", "y", "vx", "vy", "iz"):
            _same(members[j][key], o[key], "member %d %s" % (j, key))
        _same(E.get_mass(j)[0], S.get_mass()[0], "member %d mass sums" % j)
        S.close()
    E.close()


# ---- p2d-128x256-n1e6-adjoint ------------------------------------------------

def test_c5_transpose_at_workload_geometry():
    """<K rho, g> = <rho, K^T g> on the 128 x 256 grid, m <= 8 (FP64)."""
    from galaxy_amd.lib15 import P2DGrid
    w = configs.C5
    _say("C5 transposed solve")
    g = P2DGrid(precision="fp64", **configs.grid_kwargs(w))
    g.greenm()
    rng = np.random.default_rng(5)
    mfx = rng.integers(0, 1 << 40, size=(1, g.mesh), dtype=np.int64)
    g.set_mass(mfx=mfx, jrad=np.array([1], np.int32), krad=np.array([g.nr], np.int32))
    g.findf(False)
    rho, F = g.get_rho(), g.get_field()
    gg = rng.standard_normal((2, g.mesh))
    rb = g.adj_field(gg)
    lhs = float(np.dot(F[0], gg[0]) + np.dot(F[1], gg[1]))
    rhs = float(np.dot(rho, rb))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)
    g.close()


def test_c5_gradient_vs_central_differences():
    """N = 1e6, 200 steps of dt = 0.05, FP64 forward: the adjoint's
    directional derivative along a random velocity direction against
    central differences with the end state's annulus membership held fixed
    (as the adjoint holds it), at h = 1e-7 |v|; the FP32 forward's
    gradient against the FP64 one."""
    from galaxy_amd import adjoint
    w = configs.C5
    _say("C5 gradient, N = %d, %d steps" % (w.n, w.steps))
    m64 = adjoint.DiscModel(w, precision="fp64")
    g64 = m64.gradient(w.Q, w.md, w.steps, every=20, with_sel=True)
    rng = np.random.default_rng(17)
    dx, dy = rng.standard_normal(m64.n), rng.standard_normal(m64.n)
    vx, vy = m64.velocities(w.Q, w.md)
    vs = float(np.sqrt(np.mean(vx * vx + vy * vy)))
    ad = float(np.dot(g64["dvx"], dx) + np.dot(g64["dvy"], dy))
    h = 1e-7 * vs
    lp = m64.loss_v(vx + h * dx, vy + h * dy, w.md, w.steps, sel=g64["sel"])
    lm = m64.loss_v(vx - h * dx, vy - h * dy, w.md, w.steps, sel=g64["sel"])
    fd = (lp - lm) / (2 * h)
    _say("C5 adjoint %.9e, central difference %.9e, relative %.2e"
         % (ad, fd, abs(ad - fd) / abs(fd)))
    assert abs(ad - fd) <= 1e-4 * abs(fd)
    m64.close()
    m32 = adjoint.DiscModel(w, precision="fp32")
    g32 = m32.gradient(w.Q, w.md, w.steps, every=20)
    v32 = np.concatenate([g32["dvx"], g32["dvy"]])
    v64 = np.concatenate([g64["dvx"], g64["dvy"]])
    err = float(np.linalg.norm(v32 - v64) / np.linalg.norm(v64))
    _say("C5 fp32 vs fp64 velocity gradient %.2e, dQ %.6e vs %.6e"
         % (err, g32["dQ"], g64["dQ"]))
    assert err < 1e-2
    assert abs(g32["dQ"] - g64["dQ"]) <= 1e-2 * abs(g64["dQ"])
    m32.close()
