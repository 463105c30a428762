"""GPU parity of the p3d (3-D polar grid) step against the CPU oracle,
through the C ABI.

Bars, as for the p2d grid:
  - cells, zones, ring ranges, per-cell counts: bit-exact;
  - FP32 weights, fixed-point mass sums, the push (given the same field),
    time-centring and interpolation: bit-exact;
  - FP64: the two libms' log/atan2 differ by an ulp: 1e-12 .. 1e-13;
  - FP64 field solve: 1e-10 relative to the FP64 oracle; FP32 solve 1e-5;
  - the reference's own field test (EXAMPLES/p3d/ftest/run999.dat) through
    the GPU solve, against the values run999.lis:27-30 records.
"""
import numpy as np
import pytest

from galaxy_amd import configs
from tests.test_gpu_parity import _close, _same, _say
from tests.test_oracle_p3d import FT, FT_LIS, FT_POS

pytestmark = pytest.mark.gpu

PRECS = [("fp32", "f"), ("fp64", "d")]

# small two-zone model: 24 x 64 x 16 grid (outer cell / softening 3.1),
# planes 0.1 R_d apart, zone boundary at spherical radius 1.5 R_d
P3SMALL = configs.Workload3("p3d-24x64x16-test", nr=24, na=64, hash=4, n=20_000,
                            steps=6, dt=0.05, ngz=16, dz_model=0.1, eps_model=0.2,
                            nstep=(1, 2), zbr_model=(0.0, 1.5), seed=5)


@pytest.fixture(scope="module")
def o3():
    from oracle import oracle3
    return oracle3


def models(o3, prec, op, w=P3SMALL, n=None, scale=1.0, fixed=True, **over):
    """Matching oracle and GPU models with the same particles."""
    from galaxy_amd.lib15 import P3DGrid
    kw = configs.grid_kwargs3(w, n)
    kw.update(over)
    g = P3DGrid(precision=prec, **kw)
    g.greenm()
    m = o3.P3D(kw["nr"], kw["na"], kw["ngz"], kw["dzg"], kw["hash_"], kw["softl"],
               tsoft=kw["tsoft"], nsect=kw["nsect"], skip=kw.get("skip", ()),
               nzones=kw["nzones"], nstep=kw["nstep"], zbr=kw["zbr"], pmass=kw["pmass"],
               offanal=kw["offanal"], prec=op, green_tables=g.get_green(),
               jmass=kw.get("jmass", 2))
    if fixed:
        rmax = float(np.hypot(g.rgrid, g.zm))
        ff = configs.fixforce(w, rmax)
        g.set_fixforce(*ff)
        m.fixforce = ff
    x, y, z, vx, vy, vz, iz = configs.particles3(w, n, None, g.rgrid)
    if scale != 1.0:
        x, y, z = x * scale, y * scale, z * scale
        R = np.hypot(x, y)
        iz = np.where((R > g.rgrid) | (np.abs(z) > g.zm), w.nzones + 1,
                      np.minimum(iz, w.nzones)).astype(np.int32)
    g.restrt(x, y, z, vx, vy, vz, iz)
    m.load(x, y, z, vx, vy, vz, iz)
    assert float(m.zm) == g.zm and float(m.rgrid) == g.rgrid
    return m, g


def edge_positions3(g, rng, n=20000):
    """Random positions in the grid's volume plus edge cases: angles just
    below zero (t rounds up to na), the axis, exact nodes, the outermost
    ring and the top and bottom planes."""
    a = float(g.alpha)
    u = rng.uniform(0.0, g.nr - 1.0, n)
    r = np.exp(a * u) - 1.0
    ph = rng.uniform(-np.pi, np.pi, n)
    z = rng.uniform(-g.zm, g.zm, n)
    xs, ys, zs = [r * np.cos(ph)], [r * np.sin(ph)], [z]
    k = np.arange(1, 100)
    rr = np.exp(a * rng.uniform(0.5, g.nr - 1.01, len(k))) - 1.0
    zz = rng.uniform(-g.zm, g.zm, len(k))
    xs += [rr, rr, np.zeros_like(rr), -rr]
    ys += [-1e-9 * k, np.zeros_like(rr), np.zeros_like(rr), np.zeros_like(rr)]
    zs += [zz, zz, zz, zz]
    iu = rng.integers(0, g.nr - 2, 100)
    it = rng.integers(0, g.na, 100)
    ip = rng.integers(0, g.ngz, 100)
    rn = np.exp(a * iu) - 1.0
    th = 2 * np.pi * it / g.na
    xs.append(rn * np.cos(th))
    ys.append(rn * np.sin(th))
    zs.append(-g.zm + g.dzg * ip)
    ro = float(g.rgrid) * (1 - 1e-6 * rng.random(100))
    ph2 = rng.uniform(-np.pi, np.pi, 100)
    xs.append(ro * np.cos(ph2))
    ys.append(ro * np.sin(ph2))
    zs.append(np.where(rng.random(100) < 0.5, g.zm, -g.zm))
    return np.concatenate(xs), np.concatenate(ys), np.concatenate(zs)


@pytest.mark.parametrize("prec,op", PRECS)
@pytest.mark.parametrize("old", [True, False])
@pytest.mark.parametrize("jmass", [2, 3])
def test_weight3(o3, prec, op, old, jmass):
    _say("p3d weight", prec, "old" if old else "new", "jmass", jmass)
    m, g = models(o3, prec, op, n=100, nsect=2, jmass=jmass)
    dt = g.dtype
    x, y, z = (a.astype(dt) for a in edge_positions3(g, np.random.default_rng(7)))
    cell, wt = g.weight(x, y, z, old=old)
    oc, ow, _ = m.weight(x, y, z, fold=0 if old else 1)
    if prec == "fp32":
        _same(cell, oc, "cell")
        _same(wt, ow, "weights")
    else:
        u = np.log1p(np.hypot(x, y)) / float(g.alpha)
        t = np.mod(np.arctan2(y, x) * g.na / (2 * np.pi), g.na)
        tie = (np.abs(u - np.round(u)) < 1e-9) | (np.abs(t - np.round(t)) < 1e-9)
        bad = cell != oc
        assert not (bad & ~tie).any(), np.argwhere(bad & ~tie)[:5]
        _close(wt[~tie], ow[~tie], 1e-12, "weights (non-tie)")
        cell, oc = cell[~tie], oc[~tie]
    _same(np.bincount(cell, minlength=g.mesh + 2 * g.na),
          np.bincount(oc, minlength=g.mesh + 2 * g.na), "counts")
    assert wt.shape[1] == (27 if jmass == 3 else 8)


@pytest.mark.parametrize("tsoft", [1, 2])
def test_green3_fp64(o3, tsoft):
    _say("p3d greenm fp64 tsoft", tsoft)
    from galaxy_amd.lib15 import P3DGrid
    kw = configs.grid_kwargs3(P3SMALL, 100)
    kw.update(tsoft=tsoft, skip=(2,))
    g = P3DGrid(precision="fp64", **kw)
    g.greenm()
    G = g.get_green()
    lg = o3.filter_flags(g.na, g.ng, 1, (2,))
    Go, _ = o3.green3(g.nr, g.na, g.ngz, g.ng, lg, kw["dzg"], kw["softl"], tsoft)
    assert G.shape == Go.shape
    for t in range(4):
        sc = np.abs(Go[t]).max()
        assert np.abs(G[t] - Go[t]).max() <= 1e-12 * sc, t


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_reference_ftest_on_gpu(o3, prec):
    """run999: unit test masses at grid nodes in three zones (p3ckst.f:51-70)
    through set_mass, a full solve with the potential, and the values the
    reference's listing records (F_R, F_phi from the GPU's Cartesian field
    rotated back at the spoke)."""
    _say("p3d run999 ftest", prec)
    from galaxy_amd.lib15 import P3DGrid
    p = FT
    nr, na, ngz = p["nr"], p["na"], p["ngz"]
    g = P3DGrid(nr, na, ngz, p["dzg"], p["hash_"], p["softl"], tsoft=p["tsoft"],
                nzones=3, nstep=(1, 2, 4), zbr=(0.0, 1.0, 2.0), pmass=1.0, precision=prec)
    g.greenm()
    _, _, _, _, qb = g.get_mass()
    mfx = np.zeros((3, g.mesh), np.int64)
    jr = np.full(3, nr, np.int32)
    kr = np.zeros(3, np.int32)
    for r, th, z in FT_POS:
        r, th, z = min(max(r, 1), nr), min(max(th, 1), na), min(max(z, 1), ngz)
        izone = min(r // 5 + 1, 3)
        mfx[izone - 1, ((r - 1) * ngz + z - 1) * na + th - 1] += 1 << qb
        jr[izone - 1] = min(jr[izone - 1], r)
        kr[izone - 1] = max(kr[izone - 1], r)
    g.set_mass(mfx, np.zeros_like(mfx), jr, kr)
    g.set_step(0, np.zeros((2, 3), np.int32))
    g.findf(lpot=True)
    f = g.get_field().astype(np.float64)
    tol = 1e-6 if prec == "fp64" else 1e-5
    for ty, (r, z, t), v_fft, v_dir in FT_LIS:
        i = ((r - 1) * ngz + z - 1) * na + t - 1
        th = 2 * np.pi * (t - 1) / na
        v = {0: f[0][i] * np.cos(th) + f[1][i] * np.sin(th),
             1: -f[0][i] * np.sin(th) + f[1][i] * np.cos(th),
             2: f[2][i], 3: f[3][i]}[ty]
        assert abs(v - v_dir) <= tol, (ty, v, v_dir)
    # and everywhere against the oracle's solve of the same masses
    lg = o3.filter_flags(na, g.ng)
    rho = mfx.sum(0).astype(np.float64) * 2.0 ** -qb
    fo = o3.fndf3(nr, na, ngz, g.ng, lg, g.get_green(), rho, 3, 15, 1, nr, 4)
    o3._lib.orc_polcat(nr * ngz, na, fo.reshape(-1))
    for t in range(4):
        sc = np.abs(fo[t]).max()
        assert np.abs(f[t] - fo[t]).max() <= (1e-10 if prec == "fp64" else 1e-5) * sc, t


@pytest.mark.parametrize("prec,op", PRECS)
def test_masset_and_solve(o3, prec, op):
    _say("p3d masset + findf", prec)
    m, g = models(o3, prec, op)
    m.masset()
    g.masset()
    mfx, moz, jr, kr, qb = g.get_mass()
    assert qb == m.qbits
    _same(jr, m.jrad, "jrad")
    _same(kr, m.krad, "krad")
    if prec == "fp32":
        _same(mfx, m.mfx, "mass sums")
    else:
        _close(mfx * 2.0**-qb, m.mfx * 2.0**-qb, 1e-12, "mass sums")
    for lpot in (False, True):
        m.findf(lpot)
        g.findf(lpot)
        assert g.ranges() == m.ranges
        if prec == "fp32":
            _same(g.get_rho(), m.rho.astype(np.float32), "rho")
        else:
            _close(g.get_rho(), m.rho, 1e-12 * np.abs(m.rho).max(), "rho")
        f = g.get_field().astype(np.float64)
        tol = 1e-10 if prec == "fp64" else 1e-5
        for t in range(4 if lpot else 3):
            sc = np.abs(m.fld[t]).max()
            err = np.abs(f[t] - m.fld[t]).max() / sc
            assert err < tol, (lpot, t, err)


def chain3(m, g, nsteps, exact):
    """nsteps of findf + step on both sides, the oracle's field injected into
    the GPU each step (after checking the GPU's own solve against it)."""
    m.masset()
    g.masset()
    dt = m.dt
    vmax = max(np.abs(getattr(m, k)).max() for k in ("vx", "vy", "vz")) if len(m.x) else 0.0
    for k in range(nsteps):
        g.findf()
        m.findf()
        assert g.ranges() == m.ranges, k
        # the solved rings: on a partial step the reference's polcat also
        # turns the rings outside [jrf, krf] once more (polcat.f:22-37 loops
        # over every ring), which no moving particle reads
        jrf, krf = m.ranges[2], m.ranges[3]
        sl = slice((jrf - 1) * m.ngz * m.na, krf * m.ngz * m.na)
        f = g.get_field().astype(np.float64)
        for t in range(3):
            sc = max(np.abs(m.fld[t][sl]).max(), 1e-300)
            err = np.abs(f[t][sl] - m.fld[t][sl]).max() / sc
            assert err < (1e-5 if dt == np.float32 else 1e-10), ("solve", k, t, err)
        fld = m.fld.astype(dt)
        g.set_field(fld)
        m.step(field=fld[:3])
        g.step()
        out = g.unload()
        for key in ("x", "y", "z", "vx", "vy", "vz"):
            if exact:
                _same(out[key], getattr(m, key), "%s step %d" % (key, k))
            else:
                tol = 1e-13 * (float(g.rgrid) if key in "xyz" else vmax)
                _close(out[key], getattr(m, key), tol, "%s step %d" % (key, k))
        _same(out["iz"], m.iz, "zones step %d" % k)
        mfx, moz, jr, kr, qb = g.get_mass()
        if exact:
            _same(mfx, m.mfx, "mass sums step %d" % k)
            _same(moz[1:], m.mozfx[1:], "saved zone masses step %d" % k)
        else:
            _close(mfx * 2.0**-qb, m.mfx * 2.0**-qb, 1e-12, "mass sums step %d" % k)
        _same(jr, m.jrad, "jrad step %d" % k)
        _same(kr, m.krad, "krad step %d" % k)
        i, mz, ls = g.get_step()
        assert i == m.istep and mz == m.mzone
        assert g.noff == m.noff


@pytest.mark.parametrize("prec,op", PRECS)
def test_step_chain3(o3, prec, op):
    """Six steps through the two-zone cycle with the halo force; particles
    spread 1.3x so that some leave the grid through R and |z|."""
    _say("p3d step chain", prec)
    m, g = models(o3, prec, op, scale=1.3)
    assert (m.iz == 3).any()
    chain3(m, g, 6, prec == "fp32")


def test_step_chain3_options(o3):
    """Sectors, a skipped harmonic, cubic-spline softening, offanal off, no
    fixed force, FP32 bit-exact."""
    _say("p3d step chain options")
    m, g = models(o3, "fp32", "f", scale=1.3, fixed=False, nsect=2, skip=(0,), tsoft=2,
                  offanal=False)
    chain3(m, g, 4, True)


@pytest.mark.parametrize("prec,op", PRECS)
def test_tmcent3(o3, prec, op):
    _say("p3d tmcent", prec)
    m, g = models(o3, prec, op)
    m.masset()
    g.masset()
    m.findf()
    fld = m.fld.astype(m.dt)
    g.set_field(fld)
    g.tmcent(True)
    m.tmcent(True, field=fld[:3])
    out = g.unload()
    for key in ("vx", "vy", "vz"):
        if prec == "fp32":
            _same(out[key], getattr(m, key), key)
        else:
            _close(out[key], getattr(m, key), 1e-15, key)


def test_grdacc3(o3):
    _say("p3d grdacc fp32")
    m, g = models(o3, "fp32", "f", fixed=False)
    m.masset()
    m.findf()
    fld = m.fld.astype(np.float32)
    g.set_field(fld)
    x, y, z = (a.astype(np.float32) for a in edge_positions3(g, np.random.default_rng(9), 5000))
    acc = g.grdacc(x, y, z)
    m.load(x, y, z, 0 * x, 0 * x, 0 * x, np.ones(len(x), np.int32))
    _same(acc, m.getacc(field=fld[:3]), "grid accelerations")


def test_run_equals_step_loop():
    """galhip_run(k) is k x (findf; step) on the device, bit for bit."""
    _say("p3d run vs loop")
    from galaxy_amd.lib15 import P3DGrid
    kw = configs.grid_kwargs3(P3SMALL)
    parts = configs.particles3(P3SMALL)
    out = []
    for mode in ("run", "loop"):
        g = P3DGrid(precision="fp32", **kw)
        g.greenm()
        g.restrt(*parts)
        g.masset()
        if mode == "run":
            g.run(4)
        else:
            for _ in range(4):
                g.findf()
                g.step()
        out.append(g.unload())
    for key in out[0]:
        _same(out[0][key], out[1][key], key)


@pytest.mark.parametrize("prec,op", PRECS)
def test_integrals3_match_oracle(o3, prec, op):
    """galhip_integrals on the 3-D grid against the oracle's numpy sums:
    mass, centre of mass, momentum, kinetic energy and Lz; the grid
    potential energy to the solve's accuracy; the fixed-potential energy of
    the spherical radius from the same table; NaN without a potential."""
    from galaxy_amd import disc
    _say("p3d integrals", prec)
    m, g = models(o3, prec, op)
    rmax = float(np.hypot(g.rgrid, g.zm))
    fp = disc.halo_potential_table(rmax, P3SMALL.lscale(), P3SMALL.dt)
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
    tol_kin = 1e-10 if prec == "fp64" else 1e-6
    assert np.abs(vec[:9] - ref[:9]).max() <= tol_kin * scale, (vec[:9], ref[:9])
    assert abs(vec[9] - ref[9]) <= (1e-9 if prec == "fp64" else 2e-5) * abs(ref[9])
    assert abs(vec[10] - ref[10]) <= (1e-10 if prec == "fp64" else 1e-6) * abs(ref[10])


@pytest.mark.parametrize("prec,op", PRECS)
def test_step_chain3_quadratic(o3, prec, op):
    """Quadratic interpolation (jmass 3, 27 cells: weight.f:251-300,
    massgn.f:172-205, grdacc.f:259-289): the six-step chain through the
    two-zone cycle, FP32 bit-exact, FP64 to 1e-13."""
    _say("p3d step chain jmass 3", prec)
    m, g = models(o3, prec, op, scale=1.3, jmass=3)
    chain3(m, g, 6, prec == "fp32")


def test_tmcent_integrals_quadratic(o3):
    """tmcent and the integrals with jmass 3 (FP32 tmcent bit-exact)."""
    from galaxy_amd import disc
    _say("p3d tmcent + integrals jmass 3")
    m, g = models(o3, "fp32", "f", jmass=3)
    fp = disc.halo_potential_table(float(np.hypot(g.rgrid, g.zm)), P3SMALL.lscale(), P3SMALL.dt)
    g.set_fixpot(*fp)
    m.masset()
    g.masset()
    m.findf(lpot=True)
    fld = m.fld.astype(np.float32)
    g.set_field(fld)
    g.tmcent(True)
    m.tmcent(True, field=fld[:3])
    out = g.unload()
    for key in ("vx", "vy", "vz"):
        _same(out[key], getattr(m, key), key)
    g.findf(True)
    m.findf(lpot=True)
    ref = m.integrals(fixpot=fp, centred=True)
    got = g.integrals(centred=True)
    assert abs(got["w_grid"] - ref[9]) <= 2e-5 * abs(ref[9])
    assert abs(2 * got["ke"] - ref[7]) <= 1e-6 * abs(ref[7])
