"""Pins for the p3d oracle (oracle/orc_p3d.c) on the CPU.

- The reference's own p3d field test: EXAMPLES/p3d/ftest/run999.dat (grid
  22 x 64 x 45, dz = 0.25, HASH 11, cubic-spline softening 1.0, five unit
  test masses placed as p3ckst.f:51-70 does) with the values that
  run999.lis:26-30 records for each field type (FFT and direct, FP32).
- Without harmonic filtering the solve equals a brute-force sum of the
  Plummer field of every source cell (no periodic images in z: the doubling
  of p3manl removes them), written here from the closed forms.
- Trilinear weights: partition of unity and exact interpolation of linear
  functions of the grid coordinates.
- massgn totals, ranges; free drift and the off-grid rule |z| > zm.
"""
import numpy as np
import pytest

from oracle import oracle as o2
from oracle import oracle3 as o3


# run999.dat: grid size 22 64 45, z spacing 0.25, HASH 11, softl 1.0 2
FT = dict(nr=22, na=64, ngz=45, dzg=0.25, hash_=11, softl=1.0, tsoft=2)
FT_POS = [(15, 99, 3), (8, 13, 1), (3, 1, 23), (7, 2, 23), (5, 9, 45)]
# run999.lis:27-30: type, (r, z, t) of the largest error, FFT value, direct value
FT_LIS = [(0, (11, 23, 2), -1.108791590, -1.108790636),
          (1, (4, 1, 1), 0.636240482, 0.636239529),
          (2, (2, 22, 50), 0.466116756, 0.466117293),
          (3, (8, 23, 8), -2.806601524, -2.806598663)]


def ftest_rho(nr, na, ngz, pos):
    """Unit masses at grid nodes, clamped as p3ckst.f:55-61 does."""
    rho = np.zeros(nr * ngz * na)
    rmin, rmax = nr, 0
    for r, th, z in pos:
        r = min(max(r, 1), nr)
        th = min(max(th, 1), na)
        z = min(max(z, 1), ngz)
        rho[((r - 1) * ngz + z - 1) * na + th - 1] += 1.0
        rmin, rmax = min(rmin, r), max(rmax, r)
    return rho, rmin, rmax


@pytest.fixture(scope="module")
def ftest():
    p = FT
    ng = min(p["hash_"] + 1, p["na"] // 2 + 1)
    lg = o3.filter_flags(p["na"], ng)
    G, D = o3.green3(p["nr"], p["na"], p["ngz"], ng, lg, p["dzg"], p["softl"],
                     p["tsoft"], direct=True)
    rho, jrs, krs = ftest_rho(p["nr"], p["na"], p["ngz"], FT_POS)
    fld = o3.fndf3(p["nr"], p["na"], p["ngz"], ng, lg, G, rho, jrs, krs, 1, p["nr"], 4)
    return p, ng, lg, G, D, rho, fld


def test_ftest_fft_equals_direct(ftest):
    p, ng, lg, G, D, rho, fld = ftest
    for ty in range(4):
        d = o3.direct_field3(p["nr"], p["na"], p["ngz"], ty, D, rho)
        sc = np.max(np.abs(d))
        assert np.max(np.abs(fld[ty] - d)) <= 1e-12 * sc, ty


def test_ftest_matches_reference_listing(ftest):
    """run999.lis:27-30: the oracle (FP64) agrees with the reference's FP32
    FFT and direct values to their own FP32 disagreement (<= 3e-6)."""
    p, ng, lg, G, D, rho, fld = ftest
    for ty, (r, z, t), v_fft, v_dir in FT_LIS:
        i = ((r - 1) * p["ngz"] + z - 1) * p["na"] + t - 1
        v = fld[ty][i]
        assert abs(v - v_dir) <= 1e-6, (ty, v, v_dir)
        assert abs(v - v_fft) <= 3e-6, (ty, v, v_fft)
    # the recorded point is where each type's FFT-direct gap is largest,
    # so a planted sign or index error elsewhere would not show there:
    # the field must also carry the reference's symmetries
    nr, na, ngz = p["nr"], p["na"], p["ngz"]
    f = fld.reshape(4, nr, ngz, na)
    # a single source at plane 23 of 45 (the mid-plane) in ring 3 spoke 1:
    # vertical force antisymmetric, potential symmetric about it
    rho1, j, k = ftest_rho(nr, na, ngz, [(3, 1, 23)])
    f1 = o3.fndf3(nr, na, ngz, ng, lg, G, rho1, j, k, 1, nr, 4).reshape(4, nr, ngz, na)
    assert np.allclose(f1[2, :, 22 + 5, :], -f1[2, :, 22 - 5, :], atol=1e-12)
    assert np.allclose(f1[3, :, 22 + 5, :], f1[3, :, 22 - 5, :], atol=1e-12)
    # azimuthal force odd, radial even about the source spoke
    assert np.allclose(f1[1, :, 22, 5], -f1[1, :, 22, na - 5], atol=1e-12)
    assert np.allclose(f1[0, :, 22, 5], f1[0, :, 22, na - 5], atol=1e-12)
    assert f.shape == (4, nr, ngz, na)


def plummer_brute(nr, na, ngz, dzg, eps, rho):
    """Field of Plummer point masses at every source cell, by direct sums
    of the closed forms, in polar components at every cell."""
    a = 2 * np.pi / na
    R = np.exp(a * np.arange(nr)) - 1.0
    th = a * np.arange(na)
    zz = dzg * np.arange(ngz)
    Rg, Zg, Tg = np.meshgrid(R, zz, th, indexing="ij")
    X, Y = Rg * np.cos(Tg), Rg * np.sin(Tg)
    out = np.zeros((4,) + Rg.shape)
    src = np.nonzero(rho)[0]
    for s in src:
        i_s, z_s, t_s = s // (ngz * na), (s // na) % ngz, s % na
        xs, ys, zs = R[i_s] * np.cos(th[t_s]), R[i_s] * np.sin(th[t_s]), zz[z_s]
        dx, dy, dz = xs - X, ys - Y, zs - Zg
        d2 = dx * dx + dy * dy + dz * dz + eps * eps
        f = rho[s] / d2 ** 1.5
        fx, fy = f * dx, f * dy
        out[0] += fx * np.cos(Tg) + fy * np.sin(Tg)
        out[1] += -fx * np.sin(Tg) + fy * np.cos(Tg)
        out[2] += f * dz
        out[3] += -rho[s] / np.sqrt(d2)
    return out.reshape(4, -1)


def test_unfiltered_solve_is_plummer_brute_force():
    nr, na, ngz, dzg, eps = 7, 8, 6, 0.5, 0.7
    ng = na // 2 + 1
    lg = o3.filter_flags(na, ng)
    G, _ = o3.green3(nr, na, ngz, ng, lg, dzg, eps, 1)
    rng = np.random.default_rng(3)
    rho = np.zeros(nr * ngz * na)
    idx = rng.choice(rho.size, 9, replace=False)
    rho[idx] = rng.uniform(0.5, 2.0, 9)
    rings = idx // (ngz * na) + 1
    fld = o3.fndf3(nr, na, ngz, ng, lg, G, rho, rings.min(), rings.max(), 1, nr, 4)
    ref = plummer_brute(nr, na, ngz, dzg, eps, rho)
    for ty in range(4):
        sc = np.max(np.abs(ref[ty]))
        assert np.max(np.abs(fld[ty] - ref[ty])) <= 1e-11 * sc, ty


def test_partial_field_range_keeps_other_rings():
    nr, na, ngz = 6, 8, 4
    ng = 3
    lg = o3.filter_flags(na, ng)
    G, _ = o3.green3(nr, na, ngz, ng, lg, 0.5, 0.8, 1)
    rng = np.random.default_rng(5)
    rho = rng.uniform(0, 1, nr * ngz * na)
    full = o3.fndf3(nr, na, ngz, ng, lg, G, rho, 1, nr, 1, nr, 4)
    part = np.full((4, nr * ngz * na), 7.0)
    o3.fndf3(nr, na, ngz, ng, lg, G, rho, 1, nr, 2, 4, 4, part)
    ring = np.arange(nr * ngz * na) // (ngz * na) + 1
    inside = (ring >= 2) & (ring <= 4)
    assert np.allclose(part[:, inside], full[:, inside], rtol=0, atol=1e-12)
    assert np.all(part[:, ~inside] == 7.0)


def _model(prec="d", **kw):
    args = dict(nr=12, na=16, ngz=8, dzg=0.5, hash_=4, softl=0.8, tsoft=1, prec=prec)
    args.update(kw)
    return o3.P3D(**args)


def test_trilinear_weights():
    m = _model()
    rng = np.random.default_rng(7)
    n = 2000
    R = rng.uniform(0.1, 0.9 * m.rgrid, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    z = rng.uniform(-0.9 * m.zm, 0.9 * m.zm, n)
    x, y = R * np.cos(ph), R * np.sin(ph)
    cell, w, rr = m.weight(x, y, z)
    assert np.allclose(w.sum(1), 1.0, atol=1e-14)
    assert np.all(w >= -1e-15)
    # a linear function of (u, t, zg) at the corners interpolates exactly
    na, ngz = m.na, m.ngz
    u = np.log1p(rr) / m.alpha
    t = np.mod(np.arctan2(y, x) * na / (2 * np.pi), na)
    zg = (z + m.zm) / m.dzg
    ju, rem = cell // (ngz * na), cell % (ngz * na)
    jz, jt = rem // na, rem % na
    cu = ju[:, None] + np.array([0, 0, 0, 0, 1, 1, 1, 1])
    ct = jt[:, None] + np.array([0, 1, 0, 1, 0, 1, 0, 1])
    cz = jz[:, None] + np.array([0, 0, 1, 1, 0, 0, 1, 1])
    f = lambda a, b, c: 1.5 + 2.0 * a - 0.75 * b + 3.25 * c
    interp = (w * f(cu, ct, cz)).sum(1)
    assert np.allclose(interp, f(u, t, zg), atol=1e-10)


def test_weight_fp32_matches_fp64_cells():
    mf, md = _model("f"), _model("d")
    rng = np.random.default_rng(11)
    n = 5000
    R = rng.uniform(0.05, 0.95 * md.rgrid, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    z = rng.uniform(-0.95 * md.zm, 0.95 * md.zm, n)
    x, y = (R * np.cos(ph)).astype(np.float32), (R * np.sin(ph)).astype(np.float32)
    z = z.astype(np.float32)
    cf, wf, _ = mf.weight(x, y, z)
    cd, wd, _ = md.weight(x, y, z)
    same = cf == cd
    assert same.mean() > 0.999
    assert np.max(np.abs(wf[same] - wd[same])) < 2e-5


def test_massgn_total_and_ranges():
    m = _model(nzones=2, nstep=(1, 2), zbr=(0.0, 3.0))
    rng = np.random.default_rng(2)
    n = 3000
    R = rng.uniform(0.1, 0.9 * m.rgrid, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    z = rng.uniform(-0.9 * m.zm, 0.9 * m.zm, n)
    x, y = R * np.cos(ph), R * np.sin(ph)
    iz = np.where(np.sqrt(R * R + z * z) >= 3.0, 2, 1).astype(np.int32)
    m.load(x, y, z, 0 * x, 0 * x, 0 * x, iz)
    m.masset()
    q = 2.0 ** -m.qbits
    for zone in (1, 2):
        tot = m.mfx[zone - 1].sum() * q
        assert abs(tot - np.sum(iz == zone)) < 1e-9
        assert abs(m.grdmss[zone - 1].sum() - np.sum(iz == zone)) < 1e-9
        cell, _, _ = m.weight(x[iz == zone], y[iz == zone], z[iz == zone])
        ring = cell // (m.ngz * m.na)
        assert m.jrad[zone - 1] == ring.min() + 1
        assert m.krad[zone - 1] == ring.max() + 2
    # combined, scaled mass (zoncmb + scaled) equals the sum of both zones
    assert np.allclose(m.rho, (m.mfx[0] + m.mfx[1]) * q * m.pmass, atol=1e-12)


def test_free_drift_and_offgrid_rule():
    m = _model(prec="f", nzones=1)
    x = np.array([1.0, 2.0, 0.5], np.float32)
    y = np.array([0.5, -1.0, 0.2], np.float32)
    z = np.array([0.1, 0.3, m.zm - 0.05], np.float32)
    vx = np.array([0.01, 0.0, 0.0], np.float32)
    vy = np.array([0.0, 0.02, 0.0], np.float32)
    vz = np.array([0.0, 0.0, 0.1], np.float32)
    m.load(x, y, z, vx, vy, vz)
    zero = np.zeros((3, m.mesh), np.float32)
    dn = m.step(field=zero)
    assert np.array_equal(m.x, x + vx) and np.array_equal(m.y, y + vy)
    assert np.array_equal(m.z, z + vz)
    # the third particle left through the top plane: |z| > zm
    assert dn == 1 and list(m.iz) == [1, 1, 2]


def test_spherical_zone_boundaries():
    """A particle near the axis but high above the plane sits in the outer
    zone: p3d zones use spherical radii (setrun.f:810, zoneno.f:40-41)."""
    m = _model(prec="d", nzones=2, nstep=(1, 2), zbr=(0.0, 1.5))
    x = np.array([0.3, 0.3]); y = np.array([0.0, 0.0]); z = np.array([0.2, 1.6])
    m.load(x, y, z, 0 * x, 0 * x, 0 * x, np.array([1, 1], np.int32))
    m.step(field=np.zeros((3, m.mesh)))
    assert list(m.iz) == [1, 2]


def test_push_fp32_tracks_fp64():
    kw = dict(nzones=1)
    mf, md = _model("f", **kw), _model("d", **kw)
    rng = np.random.default_rng(4)
    n = 4000
    R = rng.uniform(0.5, 0.6 * md.rgrid, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    z = rng.normal(0, 0.3, n)
    x, y = R * np.cos(ph), R * np.sin(ph)
    vc = 0.02
    vx, vy = -vc * np.sin(ph), vc * np.cos(ph)
    for m in (mf, md):
        m.pmass = 1.0 / n
        m.load(x, y, z, vx, vy, 0 * z)
        m.masset()
    for _ in range(5):
        for m in (mf, md):
            m.findf()
            m.step()
    assert np.max(np.abs(mf.x - md.x)) < 1e-4
    assert np.max(np.abs(mf.z - md.z)) < 1e-4


def test_quadratic_weights():
    """jmass 3 (weight.f:251-300): the 27 weights are a partition of unity,
    non-negative, interpolate linear functions of (u, t, zg) exactly about
    the central cell, and have the second moments of the triangular-shaped
    cloud (1/4 per direction, no cross terms)."""
    m = _model(jmass=3)
    rng = np.random.default_rng(17)
    n = 3000
    R = rng.uniform(0.3, 0.9 * m.rgrid, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    z = rng.uniform(-0.9 * m.zm, 0.9 * m.zm, n)
    x, y = R * np.cos(ph), R * np.sin(ph)
    cell, w, rr = m.weight(x, y, z)
    assert w.shape == (n, 27)
    assert np.allclose(w.sum(1), 1.0, atol=1e-14)
    assert np.all(w >= -1e-15)
    na, ngz = m.na, m.ngz
    u = np.log1p(rr) / m.alpha
    t = np.mod(np.arctan2(y, x) * na / (2 * np.pi), na)
    zg = (z + m.zm) / m.dzg
    ju, rem = cell // (ngz * na), cell % (ngz * na)
    jz, jt = rem // na, rem % na
    off = np.array([-1, 0, 1])
    ct = jt[:, None] + np.tile(off, 9)[None, :]
    cz = jz[:, None] + np.tile(np.repeat(off, 3), 3)[None, :]
    cu = ju[:, None] + np.repeat(off, 9)[None, :]
    # unwrap the angle of cells near spoke 0 against t
    ct = np.where(ct - t[:, None] > na / 2, ct - na, ct)
    ct = np.where(t[:, None] - ct > na / 2, ct + na, ct)
    f = lambda a, b, c: 0.5 - 1.25 * a + 0.5 * b + 2.0 * c
    assert np.allclose((w * f(cu, ct, cz)).sum(1), f(u, t, zg), atol=1e-9)
    # second moments about the particle: 1/4 in each direction, no cross
    # terms -- the triangular-shaped cloud (three unit boxes convolved).
    # With the zeroth and first moments this fixes the three 1-D weights
    # of each direction, so a wrong coefficient of the quadratic fails here.
    du, dt, dz = cu - u[:, None], ct - t[:, None], cz - zg[:, None]
    for a, b, want in ((du, du, 0.25), (dt, dt, 0.25), (dz, dz, 0.25),
                       (du, dt, 0.0), (du, dz, 0.0), (dt, dz, 0.0)):
        assert np.allclose((w * a * b).sum(1), want, atol=1e-9)


def test_quadratic_massgn_and_geometry():
    """jmass 3: rgrid = grofu(nr - 1.5), zm = dzg (ngz - 2)/2
    (setgrd.f:135-150); ring ranges j..j+2 about the central ring
    (massgn.f:176-179); mass conserved."""
    m = _model(jmass=3, prec="d")
    assert abs(m.rgrid - (np.exp(m.alpha * (m.nr - 1.5)) - 1.0)) < 1e-12
    assert abs(m.zm - 0.5 * m.dzg * (m.ngz - 2)) < 1e-15
    rng = np.random.default_rng(21)
    n = 2000
    R = rng.uniform(0.5, 0.9 * m.rgrid, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    z = rng.uniform(-0.8 * m.zm, 0.8 * m.zm, n)
    x, y = R * np.cos(ph), R * np.sin(ph)
    m.load(x, y, z, 0 * x, 0 * x, 0 * x)
    m.masset()
    assert abs(m.mfx[0].sum() * 2.0 ** -m.qbits - n) < 1e-9
    cell, _, _ = m.weight(x, y, z)
    ring = cell // (m.ngz * m.na)
    assert m.jrad[0] == ring.min() and m.krad[0] == ring.max() + 2


def _ring_particles(m, n, rng, zones=True):
    """Particles inside the grid's volume, a few off it, two zones."""
    R = rng.uniform(0.3, 0.85 * m.rgrid, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    z = rng.uniform(-0.8 * m.zm, 0.8 * m.zm, n)
    x, y = R * np.cos(ph), R * np.sin(ph)
    vx, vy, vz = rng.normal(0, 0.1, n), rng.normal(0, 0.1, n), rng.normal(0, 0.05, n)
    iz = np.where(np.sqrt(R * R + z * z) >= m.zbr[1], 2, 1) if zones else np.ones(n, int)
    iz = np.asarray(iz, np.int32)
    iz[:5] = m.nzones + 1
    x[:5] *= 5.0
    return x, y, z, vx, vy, vz, iz


@pytest.mark.parametrize("jmass", [2, 3])
def test_tmcent3_closed_form(jmass):
    """orc3_tmcent with a uniform field F0 and a constant radial fixed force
    c (a table of a constant, which cubic Hermite reproduces exactly): every
    on-grid particle's velocity moves by -(F0 + c x / r) s^2 / 2, s the step
    factor of its zone (getacc.f:142-151, tmcgrp.f:26-36); off-grid ones do
    not move; tmcent(False) takes it back."""
    m = _model(jmass=jmass, nzones=2, nstep=(1, 2), zbr=(0.0, 3.0))
    rng = np.random.default_rng(5)
    x, y, z, vx, vy, vz, iz = _ring_particles(m, 500, rng)
    F0 = np.array([0.013, -0.027, 0.041])
    m.fld[:3] = F0[:, None]
    cval = -0.05
    m.fixforce = (0.5, np.full(200, cval), np.zeros(200))
    m.load(x, y, z, vx, vy, vz, iz)
    m.tmcent(True)
    r = np.sqrt(x * x + y * y + z * z)
    s2 = np.where(iz == 2, 4.0, 1.0)
    on = iz <= m.nzones
    for d, (v0, p) in enumerate(((vx, x), (vy, y), (vz, z))):
        want = np.where(on, v0 - 0.5 * (F0[d] + cval * p / r) * s2, v0)
        got = (m.vx, m.vy, m.vz)[d]
        assert np.abs(got - want).max() < 1e-14, d
    m.tmcent(False)
    for got, v0 in ((m.vx, vx), (m.vy, vy), (m.vz, vz)):
        assert np.abs(got - v0).max() < 1e-14


@pytest.mark.parametrize("jmass", [2, 3])
def test_integrals3_closed_form(jmass):
    """oracle3.integrals (icheck.f:24-70 on the 3-D grid) against sums
    written out here: mass, m x position, momentum, sum m v^2 and Lz with
    velocities over the zone's step factor (the off-grid list's that of the
    last zone, with which it moves); a uniform potential Phi0 gives
    sum m Phi_grid = pmass x (particles on the grid) x Phi0 (partition of
    unity of either scheme); a fixed potential a + b r tabulated on a
    Hermite table gives pmass sum (a + b r) over every particle; centred
    with a uniform field adds F0 s^2 / 2 to each on-grid velocity before the
    step factor."""
    m = _model(jmass=jmass, nzones=2, nstep=(1, 2), zbr=(0.0, 3.0))
    rng = np.random.default_rng(9)
    x, y, z, vx, vy, vz, iz = _ring_particles(m, 400, rng)
    m.pmass = 0.37
    m.load(x, y, z, vx, vy, vz, iz)
    phi0, F0 = -1.75, np.array([0.01, 0.02, -0.03])
    m.fld[3] = phi0
    m.fld[:3] = F0[:, None]
    a, b, dr = 0.4, -0.3, 0.25
    rk = dr * np.arange(4000)                    # past every particle's radius
    assert rk[-1] > np.sqrt(x * x + y * y + z * z).max()
    fp = (dr, a + b * rk, np.full(4000, b))
    # step factor of the zone; the off-grid list moves with the last zone
    s = np.where(iz >= 2, 2.0, 1.0)
    on = iz <= m.nzones
    for centred in (False, True):
        out = m.integrals(fixpot=fp, centred=centred)
        kick = np.where(on, 0.5 * s * s, 0.0) if centred else np.zeros(len(x))
        v = [(vx + kick * F0[0]) / s, (vy + kick * F0[1]) / s, (vz + kick * F0[2]) / s]
        pm = 0.37
        want = np.array([pm * len(x), pm * x.sum(), pm * y.sum(), pm * z.sum(),
                         pm * v[0].sum(), pm * v[1].sum(), pm * v[2].sum(),
                         pm * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2).sum(),
                         pm * (x * v[1] - y * v[0]).sum(),
                         pm * on.sum() * phi0,
                         pm * (a + b * np.sqrt(x * x + y * y + z * z)).sum()])
        assert np.allclose(out, want, rtol=1e-13, atol=1e-13), (centred, out - want)
