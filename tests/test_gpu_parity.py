"""GPU parity of the p2d step against the CPU oracle, through the C ABI.

Bars (from the workload definition):
  - cell indices, per-cell particle counts, zones, ring ranges: bit-exact;
  - FP32: weights, fixed-point mass sums and the particle push given the
    same field are bit-exact (same operations in the same order, no fused
    multiply-add, correctly rounded logf on both sides);
  - FP64: the oracle's glibc log/atan2 and the device library's differ by
    an ulp, so weights, masses and the push agree to 1e-12..1e-13;
  - FP64 field solve: 1e-10 relative to the FP64 oracle (ftest-style);
  - FP32 field solve: relative force error <= 1e-5 against the FP64 oracle.
"""
import numpy as np
import pytest

from tests._model import SMALL, gpu_model, oracle_model, rel_err

pytestmark = pytest.mark.gpu

PRECS = [("fp32", "f"), ("fp64", "d")]


def _say(*a):
    print("[case]", *a, flush=True)


def _same(a, b, what):
    """Bit-exact comparison with a useful message."""
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    bad = a != b
    if bad.any():
        idx = np.argwhere(bad)[:5]
        d = np.abs(a.astype(np.float64) - b.astype(np.float64)).max()
        raise AssertionError("%s: %d of %d differ (max |diff| %.3g), first at %s: %s vs %s"
                             % (what, bad.sum(), bad.size, d, idx.tolist(),
                                a[bad][:5], b[bad][:5]))


def _close(a, b, atol, what):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.abs(a - b).max() if a.size else 0.0
    assert d <= atol, "%s: max |diff| %.3g > %.3g" % (what, d, atol)


def _edge_positions(g, rng, n=20000):
    """Random positions over the grid plus the reference's edge cases:
    angles just below zero (t rounds up to na), exactly on axes, cell
    corners, the outermost ring."""
    a = float(g.alpha)
    u = rng.uniform(0.0, g.nr - 1.0, n)
    r = np.exp(a * u) - 1.0
    ph = rng.uniform(-np.pi, np.pi, n)
    x, y = r * np.cos(ph), r * np.sin(ph)
    k = np.arange(1, 200)
    rr = np.exp(a * rng.uniform(0.5, g.nr - 1.01, len(k))) - 1.0
    ex = [rr, rr, -rr, rr, np.zeros_like(rr)]
    ey = [-1e-9 * k, np.zeros_like(rr), np.zeros_like(rr), 1e-30 * k, rr]
    # exact nodes of the grid
    iu = rng.integers(0, g.nr - 2, 100)
    it = rng.integers(0, g.na, 100)
    rn = np.exp(a * iu) - 1.0
    th = 2 * np.pi * it / g.na
    ex.append(rn * np.cos(th))
    ey.append(rn * np.sin(th))
    # the last ring, just inside the edge
    ro = float(g.rgrid) * (1 - 1e-6 * rng.random(100))
    ph2 = rng.uniform(-np.pi, np.pi, 100)
    ex.append(ro * np.cos(ph2))
    ey.append(ro * np.sin(ph2))
    ex.append(float(g.rgrid) * (1 - 1e-6) * np.ones(5))
    ey.append(-1e-7 * np.arange(1, 6))
    return np.concatenate([x] + ex), np.concatenate([y] + ey)


@pytest.mark.parametrize("prec,op", PRECS)
@pytest.mark.parametrize("old", [True, False])
def test_weight_bit_exact(orc, prec, op, old):
    _say("weight", prec, "old" if old else "new")
    g = gpu_model(precision=prec, n=1000)
    rng = np.random.default_rng(7)
    x, y = _edge_positions(g, rng)
    dt = np.float32 if prec == "fp32" else np.float64
    x, y = x.astype(dt), y.astype(dt)
    cell, wt = g.weight(x, y, old=old)
    oc, ow, _ = orc.weight(x, y, g.nr, g.na, op, fold=0, nsect=1)
    if prec == "fp32":
        _same(cell, oc, "cell")
        _same(wt, ow, "weights")
    else:
        # FP64: the two libms' log/atan2 differ by <= 1 ulp, which decides
        # the floor only at exact ties (positions on a grid node to 1e-15).
        # Tie audit: every mismatch must be a tie, and ties are only the
        # nodes placed deliberately.
        u = np.log1p(np.hypot(x, y)) / float(g.alpha)
        t = np.mod(np.arctan2(y, x) * g.na / (2 * np.pi), g.na)
        tie = (np.abs(u - np.round(u)) < 1e-9) | (np.abs(t - np.round(t)) < 1e-9)
        bad = cell != oc
        assert not (bad & ~tie).any(), np.argwhere(bad & ~tie)[:5]
        _same(cell[~tie], oc[~tie], "cell (non-tie)")
        print("FP64 tie audit: %d ties, %d resolved differently" % (tie.sum(), bad.sum()))
        _close(wt[~tie], ow[~tie], 1e-12, "weights (non-tie)")
        cell, oc = cell[~tie], oc[~tie]
    # per-cell particle counts
    _same(np.bincount(cell, minlength=g.mesh + g.na),
          np.bincount(oc, minlength=g.mesh + g.na), "counts")


def test_weight_sector_fold(orc):
    """Mass assignment folds t into the first sector (weight.f:80-81)."""
    _say("weight fold nsect=2")
    from galaxy_amd.lib15 import P2DGrid
    g = P2DGrid(64, 128, 8, 0.35, nsect=2, precision="fp32")
    rng = np.random.default_rng(3)
    x, y = _edge_positions(g, rng, 5000)
    x, y = x.astype(np.float32), y.astype(np.float32)
    cell, wt = g.weight(x, y, old=False)
    oc, ow, _ = orc.weight(x, y, 64, 128, "f", fold=1, nsect=2)
    _same(cell, oc, "cell")
    _same(wt, ow, "weights")
    assert (cell % 128).max() <= 64


def test_green_tables_fp64(orc):
    _say("greenm fp64")
    g = gpu_model(precision="fp64", n=100)
    G = g.get_green()
    Go, _ = orc.green(g.nr, g.na, g.ng, 1, orc.alpha(g.na), 0.1 * SMALL.lscale(), 1)
    assert G.shape == Go.shape
    for t in range(3):
        assert rel_err(G[t], Go[t]) < 1e-12, t


@pytest.mark.parametrize("prec,op", PRECS)
def test_masset_bit_exact(orc, prec, op):
    _say("masset", prec)
    m, parts = oracle_model(orc, prec=op)
    g = gpu_model(precision=prec, parts=parts)
    m.masset()
    g.masset()
    mfx, moz, jr, kr, qb = g.get_mass()
    assert qb == m.qbits
    _same(jr, m.jrad, "jrad")
    _same(kr, m.krad, "krad")
    g.findf()
    if prec == "fp32":
        _same(mfx, m.mfx, "mass sums")
        _same(g.get_rho(), m.rho.astype(np.float32), "rho")
    else:
        _close(mfx * 2.0**-qb, m.mfx * 2.0**-qb, 1e-12, "mass sums")
        _close(g.get_rho(), m.rho, 1e-12 * np.abs(m.rho).max(), "rho")


@pytest.mark.parametrize("prec,op", PRECS)
def test_findf_accuracy(orc, prec, op):
    _say("findf", prec)
    m, parts = oracle_model(orc, prec=op)
    g = gpu_model(precision=prec, parts=parts)
    m.masset()
    g.masset()
    for lpot in (False, True):
        m.findf(lpot)
        g.findf(lpot)
        assert g.ranges() == m.ranges
        f = g.get_field().astype(np.float64)
        tol = 1e-10 if prec == "fp64" else 1e-5
        for t in range(3 if lpot else 2):
            # relative force error over the whole grid
            assert rel_err(f[t], m.fld[t]) < tol, (lpot, t, rel_err(f[t], m.fld[t]))


@pytest.mark.parametrize("prec,op", PRECS)
def test_step_chain_bit_exact(orc, prec, op):
    """Eight steps through a full zone cycle (mzone 3,1,2,1,3,...), with the
    oracle's field injected each step: positions, velocities, zones, mass
    sums, ring ranges and combined masses stay bit-identical in FP32; in
    FP64 (where two libms' log/atan2 differ by an ulp) zones and ranges stay
    exact and the rest agree to 1e-13 relative."""
    _say("step chain", prec)
    m, parts = oracle_model(orc, prec=op)
    g = gpu_model(precision=prec, parts=parts)
    _chain(m, g, 8, prec == "fp32")


def _chain(m, g, nsteps, exact):
    """nsteps of findf + step on both sides with the oracle's field injected
    into the GPU each step; everything compared after every step."""
    m.masset()
    g.masset()
    dt = m.dt
    vmax = max(np.abs(m.vx).max() if len(m.vx) else 0.0,
               np.abs(m.vy).max() if len(m.vy) else 0.0)
    for k in range(nsteps):
        g.findf()                       # combined mass + ranges on the GPU
        m.findf()
        if exact:
            _same(g.get_rho(), m.rho.astype(dt), "rho step %d" % k)
        else:
            _close(g.get_rho(), m.rho, 1e-12 * np.abs(m.rho).max(), "rho step %d" % k)
        assert g.ranges() == m.ranges, k
        fld = m.fld.astype(dt)
        g.set_field(fld)
        m.step(field=fld[:2])
        g.step()
        out = g.unload()
        for key in ("x", "y", "vx", "vy"):
            if exact:
                _same(out[key], getattr(m, key), "%s step %d" % (key, k))
            else:
                tol = 1e-13 * (float(g.rgrid) if key in "xy" else vmax)
                _close(out[key], getattr(m, key), tol, "%s step %d" % (key, k))
        _same(out["iz"], m.iz, "zones step %d" % k)
        mfx, moz, jr, kr, qb = g.get_mass()
        if exact:
            _same(mfx, m.mfx, "mass sums step %d" % k)
            _same(moz[1:], m.mozfx[1:], "saved zone masses step %d" % k)
        else:
            _close(mfx * 2.0**-qb, m.mfx * 2.0**-qb, 1e-12, "mass sums step %d" % k)
            _close(moz[1:] * 2.0**-qb, m.mozfx[1:] * 2.0**-qb, 1e-12, "saved masses %d" % k)
        _same(jr, m.jrad, "jrad step %d" % k)
        _same(kr, m.krad, "krad step %d" % k)
        i, mz, ls = g.get_step()
        assert i == m.istep and mz == m.mzone
        assert np.array_equal(ls[:, 1:], m.lstep[:, 1:])
        assert g.noff == m.noff


# Grid, filter and zone configurations beyond the default model, each run
# six steps bit-exactly in FP32 (the reference's option space on the path:
# sectors and skipped harmonics, many zones, the largest supported na and
# harmonic count, particles off the grid with and without offanal, a
# single particle and none).
CONFIGS = {
    "sectors2-skip0-1zone-nooff": dict(nr=48, na=96, hash=4, nsect=2, skip=(0,),
                                       nstep=(1,), zbr=(0.0,), offanal=False),
    "4zones": dict(nr=48, na=96, hash=6, nstep=(1, 2, 4, 8), zbr=(0.0, 1.0, 2.0, 3.0)),
    "na2048": dict(nr=24, na=2048, hash=8, n=6000),
    "ng64": dict(nr=32, na=256, hash=63),
    "offgrid": dict(nr=40, na=96, hash=4, nstep=(1, 2), zbr=(0.0, 2.0), scale=1.4),
    "offgrid-nooff": dict(nr=40, na=96, hash=4, nstep=(1, 2), zbr=(0.0, 2.0), scale=1.4,
                          offanal=False),
    "one-particle": dict(nr=24, na=96, hash=2, n=1),
    "no-particles": dict(nr=24, na=96, hash=2, n=0),
}


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_config_chain_fp32(orc, name):
    from galaxy_amd import configs
    from galaxy_amd.lib15 import P2DGrid
    _say("config chain", name)
    c = dict(CONFIGS[name])
    n = c.pop("n", 4000)
    scale = c.pop("scale", 1.0)
    w = configs.Workload(name, nr=c["nr"], na=c["na"], hash=c["hash"], n=max(n, 1),
                         steps=6, dt=0.05, nstep=c.get("nstep", (1, 2, 4)),
                         zbr_model=c.get("zbr", (0.0, 1.5, 3.5)), seed=17)
    kw = configs.grid_kwargs(w, max(n, 1))
    kw.update(nsect=c.get("nsect", 1), offanal=c.get("offanal", True))
    skip = c.get("skip", ())
    m = orc.P2D(kw["nr"], kw["na"], kw["hash_"], kw["softl"], tsoft=1, nsect=kw["nsect"],
                skip=skip, nzones=kw["nzones"], nstep=kw["nstep"], zbr=kw["zbr"],
                pmass=kw["pmass"], offanal=kw["offanal"], prec="f", use_fx=True)
    m.fixforce = configs.fixforce(w, float(m.rgrid))
    x, y, vx, vy, iz = configs.particles(w, max(n, 1), None, float(m.rgrid))
    if scale != 1.0:
        x, y = x * scale, y * scale
        r = np.hypot(x, y)
        iz = np.where(r > float(m.rgrid), len(kw["nstep"]) + 1, iz).astype(np.int32)
    x, y, vx, vy, iz = (a[:n] for a in (x, y, vx, vy, iz))
    m.load(x, y, vx, vy, iz)
    g = P2DGrid(precision="fp32", skip=skip, **kw)
    g.greenm()
    g.set_fixforce(*m.fixforce)
    g.restrt(x, y, vx, vy, iz)
    if n == 0:
        # nothing to assign or move: the calls succeed and the grid stays empty
        g.masset()
        g.findf()
        g.step()
        mfx, _, _, _, _ = g.get_mass()
        assert not mfx.any() and g.noff == 0
        return
    _chain(m, g, 6, True)


def test_ftest_fp64(orc):
    """ptest/ftest (SRC/progs/ptest.f:60-88, p2chkf.f): unit masses on grid
    cells; the GPU's FFT-convolution field equals the direct convolution
    with the filtered kernel to 1e-10 relative in FP64."""
    _say("ftest fp64")
    from galaxy_amd.lib15 import P2DGrid
    nr, na, hash_, eps = 22, 64, 10, 1.0
    g = P2DGrid(nr, na, hash_, eps, precision="fp64", pmass=1.0)
    g.greenm()
    mesh = nr * na
    cells = [(8, 13), (3, 1), (7, 2), (15, 35), (5, 9)]
    _, _, _, _, qb = g.get_mass()
    mfx = np.zeros((1, mesh), np.int64)
    rho = np.zeros(mesh)
    for ir, it in cells:
        mfx[0, (ir - 1) * na + it - 1] = 1 << qb
        rho[(ir - 1) * na + it - 1] = 1.0
    g.set_mass(mfx=mfx, jrad=[1], krad=[nr])
    g.findf(lpot=True)
    f = g.get_field()
    _, D = orc.green(nr, na, hash_ + 1, 1, orc.alpha(na), eps, 1, direct=True)
    ref = np.stack([orc.direct_field(nr, na, t, D, rho) for t in range(3)])
    orc._lib.orc_polcat(nr, na, ref.reshape(-1))
    for t in range(3):
        assert rel_err(f[t], ref[t]) < 1e-10, t


def test_run_fp64_matches_oracle(orc):
    """galhip_run (fused masscl, periodic re-sort) over 8 steps against the
    FP64 oracle's own solve + step."""
    _say("run fp64")
    m, parts = oracle_model(orc, prec="d")
    g = gpu_model(precision="fp64", parts=parts)
    g.set_sort_interval(3)
    m.masset()
    g.masset()
    for _ in range(8):
        m.findf()
        m.step()
    g.run(8)
    out = g.unload()
    scale = float(g.rgrid)
    for key in ("x", "y"):
        assert np.abs(out[key] - getattr(m, key)).max() < 1e-9 * scale
    assert np.array_equal(out["iz"], m.iz)
    i, mz, _ = g.get_step()
    assert i == 8 and mz == m.mzone


def test_sort_preserves_particles():
    _say("sort")
    g = gpu_model(precision="fp32", n=20000)
    before = g.unload()
    g.sort()
    after = g.unload()
    for k in before:
        assert np.array_equal(before[k], after[k])


def test_sort_groups_by_cell(orc):
    """After galhip_sort the storage order runs through the cells in
    ring-major order, off-grid particles last (the cells recomputed by the
    oracle from the unloaded positions)."""
    _say("sort grouping")
    n = 20000
    g = gpu_model(precision="fp32", n=n)
    g.run(3)
    g.sort()
    ids = g.get_order()
    out = g.unload()
    assert np.array_equal(np.sort(ids), np.arange(n))
    nr, na, nz = SMALL.nr, SMALL.na, len(SMALL.nstep)
    cell, _, _ = orc.weight(out["x"], out["y"], nr, na, "f")
    on = (out["iz"] >= 1) & (out["iz"] <= nz)
    key = np.where(on, cell, nr * na)
    assert np.all(np.diff(key[ids]) >= 0)


@pytest.mark.parametrize("prec,op", PRECS)
def test_integrals_match_oracle(orc, prec, op):
    """galhip_integrals (icheck's global integrals, FP64 sums) against the
    oracle's numpy sums over the same particles: mass, centre of mass,
    momentum, kinetic energy and Lz; the grid potential energy to
    the solve's own accuracy; the fixed-potential energy from the same
    table.  Without a potential the grid term
    is NaN."""
    from galaxy_amd import disc
    _say("integrals", prec)
    m, parts = oracle_model(orc, prec=op)
    g = gpu_model(precision=prec, parts=parts)
    fp = disc.halo_potential_table(float(g.rgrid), SMALL.lscale(), SMALL.dt)
    g.set_fixpot(*fp)
    m.masset()
    g.masset()
    g.findf(False)
    assert np.isnan(g.integrals(centred=False)["w_grid"])
    for _ in range(3):
        m.findf()
        m.step()
    g.run(3)
    m.findf(lpot=True)
    g.findf(True)
    ref = m.integrals(fixpot=fp, centred=True)
    got = g.integrals(centred=True)
    vec = np.concatenate([[got["mass"]], got["cm"] * got["mass"], got["momentum"],
                          [2 * got["ke"], got["lz"], got["w_grid"], got["w_fixed"]]])
    scale = np.abs(ref).max()
    # the three steps ran on each side's own solve: positions agree to the
    # solves' accuracy (FP64 1e-10, FP32 1e-5 in the force), hence the bounds
    tol_kin = 1e-10 if prec == "fp64" else 1e-6
    assert np.abs(vec[:7] - ref[:7]).max() <= tol_kin * scale, (vec[:7], ref[:7])
    assert abs(vec[7] - ref[7]) <= (1e-9 if prec == "fp64" else 2e-5) * abs(ref[7])
    assert abs(vec[8] - ref[8]) <= (1e-10 if prec == "fp64" else 1e-6) * abs(ref[8])
