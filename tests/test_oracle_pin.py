"""Pins the CPU oracle against independent mathematics.

Each check is chosen so that a plausible slip in the oracle (a dropped term,
a wrong sign or index, a transposed operand) fails it:
  - FFTPACK conventions against numpy / scipy FFTs
    (SRC/utils/sfftpack/stest.f:84-195 define the conventions);
  - Green's tables against kernels rebuilt from Cartesian vector geometry
    and numpy FFTs (SRC/lib15/radfor.f, azifor.f, sftpot.f, p2grnm.f);
  - the ftest identity: FFT convolution == direct convolution with the
    filtered kernel (SRC/lib15/p2chkf.f:81-128);
  - brute-force particle-mesh field on a tiny unfiltered grid;
  - CIC weights: exact reproduction of bilinear fields, and the
    assignment/interpolation transpose pair (weight.f, massgn.f, grdacc.f);
  - the leapfrog against the closed-form discrete harmonic oscillator
    (stpgrp.f), and chstep with tfac = 1 as the identity (chstep.f).
"""
import numpy as np
import pytest
import scipy.fft as sf


def _hc(X, n):
    hc = np.empty(n)
    hc[0] = X[0].real
    for k in range(1, (n + 1) // 2):
        hc[2 * k - 1] = X[k].real
        hc[2 * k] = X[k].imag
    if n % 2 == 0:
        hc[n - 1] = X[n // 2].real
    return hc


@pytest.mark.parametrize("n", [8, 12, 64, 128, 512])
def test_rfft_conventions(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    hc = _hc(np.fft.rfft(x), n)
    assert np.allclose(orc.rfftf(x), hc, rtol=0, atol=1e-12 * n)
    # backward is unnormalised: B(F(x)) = n x
    assert np.allclose(orc.rfftb(orc.rfftf(x)), n * x, rtol=0, atol=1e-11 * n)
    # backward closed form: y_i = a0 + 2 sum(a_k cos - b_k sin) + (-1)^i a_n/2
    y = orc.rfftb(hc)
    assert np.allclose(y, n * np.fft.irfft(np.fft.rfft(x), n), atol=1e-11 * n)


@pytest.mark.parametrize("n", [5, 17, 65, 257])
def test_dct_dst(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    assert np.allclose(orc.cost(x), sf.dct(x, type=1), atol=1e-12 * n)
    assert np.allclose(orc.sint(x), sf.dst(x, type=1), atol=1e-12 * n)


def _kernel_geometry(rs, rf, th, eps, tsoft):
    """Force and potential of a unit mass at (rs, 0) seen at (rf cos th,
    rf sin th), from Cartesian vectors; returns (F_R, F_phi, Phi)."""
    sx, sy = rs, 0.0
    fx, fy = rf * np.cos(th), rf * np.sin(th)
    dx, dy = fx - sx, fy - sy
    d2 = dx * dx + dy * dy
    d = np.sqrt(d2)
    if tsoft == 1:
        g = 1.0 / (d2 + eps * eps) ** 1.5
        pot = -1.0 / np.sqrt(d2 + eps * eps)
    else:
        x = d / eps
        m = np.where(x < 1, 4 * x**3 * (1 / 3 - 0.3 * x**2 + 0.125 * x**3),
             np.where(x < 2, -1 / 15 + x**3 * (8 / 3 - 3 * x + 1.2 * x**2 - x**3 / 6), 1.0))
        with np.errstate(divide="ignore", invalid="ignore"):
            g = np.where(d2 > 0, m / d2**1.5, 0.0)
            pot = np.where(x < 1, (-1.4 + x**2 * (2 / 3 - 0.3 * x**2 + 0.1 * x**3)) / eps,
                  np.where(x < 2, (-1.6 + 1 / (15 * x) + x**2 * (4 / 3 - x + 0.3 * x**2 - x**3 / 30)) / eps,
                           -1 / d))
    ax, ay = -dx * g, -dy * g          # attraction towards the source
    fr = ax * np.cos(th) + ay * np.sin(th)
    fp = -ax * np.sin(th) + ay * np.cos(th)
    return fr, fp, pot


@pytest.mark.parametrize("tsoft", [1, 2])
def test_green_tables_vs_geometry(orc, tsoft):
    nr, na, nsect = 8, 16, 1
    mmax = na // 2 + 1
    a = orc.alpha(na)
    eps = 0.37
    ng = mmax
    G, _ = orc.green(nr, na, ng, nsect, a, eps, tsoft)
    th = 2 * np.pi * np.arange(na) / na
    for is_ in range(nr):
        rs = np.exp(a * is_) - 1
        for ir in range(nr):
            rf = np.exp(a * ir) - 1
            fr, fp, pot = _kernel_geometry(rs, rf, th, eps, tsoft)
            Fr, Fp, Fpot = np.fft.fft(fr), np.fft.fft(fp), np.fft.fft(pot)
            for m in range(ng):
                scale = 1 + abs(Fr).max()
                assert abs(G[0, m, ir, is_] - Fr[m].real) < 1e-11 * scale
                assert abs(G[2, m, ir, is_] - Fpot[m].real) < 1e-11 * (1 + abs(Fpot).max())
                if 1 <= m <= mmax - 2:
                    # odd kernel: table = sum F sin(m th) = -Im FFT
                    assert abs(G[1, m, ir, is_] + Fp[m].imag) < 1e-11 * (1 + abs(Fp).max())


def test_green_filter(orc):
    nr, na = 6, 32
    a = orc.alpha(na)
    G, _ = orc.green(nr, na, 7, 2, a, 0.5, 1)
    assert G.shape == (3, 7, nr, nr)
    for m in (1, 3, 5):
        assert np.all(G[:, m] == 0)        # m mod nsect != 0 removed
    assert np.all(G[1, 0] == 0)            # no m = 0 azimuthal table
    assert np.any(G[0, 2] != 0) and np.any(G[1, 4] != 0)


def _solve(orc, nr, na, ng, nsect, eps, tsoft, rho, lg=None, types=3):
    a = orc.alpha(na)
    G, D = orc.green(nr, na, ng, nsect, a, eps, tsoft, direct=True)
    mmax = na // 2 + 1
    if lg is None:
        lg = np.zeros(mmax, np.int32)
        for m in range(1, ng):
            lg[m] = 0 if m % nsect == 0 else 1
    fld = np.zeros((3, nr * na))
    r = rho.copy()
    orc._lib.orc_fndf(nr, na, ng, lg, G.reshape(-1), r, 1, nr, 1, nr, types,
                      fld.reshape(-1))
    return fld, G, D


@pytest.mark.parametrize("ng,nsect,tsoft", [(11, 1, 1), (11, 2, 1), (33, 1, 1), (8, 1, 2)])
def test_ftest_identity(orc, ng, nsect, tsoft):
    """FFT convolution equals direct convolution with the filtered kernel,
    to 1e-10 relative in FP64 (the reference reaches ~1e-6 in FP32,
    EXAMPLES/p3d/ftest/run999.lis:26-30)."""
    nr, na = 22, 64
    rho = np.zeros(nr * na)
    # unit masses at cells in the manner of EXAMPLES/p2d/ftest/run999.dat
    for ir, it in ((8, 13), (3, 1), (7, 2), (15, 35), (5, 9)):
        rho[(ir - 1) * na + (it - 1)] = 1.0
    fld, G, D = _solve(orc, nr, na, ng, nsect, 1.0, tsoft, rho)
    for t in range(3):
        ref = orc.direct_field(nr, na, t, D, rho)
        scale = np.abs(ref).max()
        err = np.abs(fld[t] - ref).max() / scale
        assert err < 1e-10, (t, err)


def test_brute_force_pm_tiny(orc):
    """All harmonics active: the solved field equals the brute-force sum of
    point-mass kernels over every source cell."""
    nr, na, eps = 8, 16, 0.6
    a = orc.alpha(na)
    rng = np.random.default_rng(3)
    rho = rng.random(nr * na)
    fld, _, _ = _solve(orc, nr, na, na // 2 + 1, 1, eps, 1, rho)
    r = np.exp(a * np.arange(nr)) - 1
    th = 2 * np.pi * np.arange(na) / na
    ref = np.zeros((3, nr * na))
    for s in range(nr * na):
        is_, js = divmod(s, na)
        for ir in range(nr):
            fr, fp, pot = _kernel_geometry(r[is_], r[ir], th - th[js], eps, 1)
            ref[0, ir * na:(ir + 1) * na] += rho[s] * fr
            ref[1, ir * na:(ir + 1) * na] += rho[s] * fp
            ref[2, ir * na:(ir + 1) * na] += rho[s] * pot
    for t in range(3):
        assert np.abs(fld[t] - ref[t]).max() < 1e-11 * np.abs(ref[t]).max()


def test_polcat_rotation(orc):
    nr, na = 3, 16
    f = np.zeros((3, nr * na))
    f[0] = 1.0                                   # unit radial field
    orc._lib.orc_polcat(nr, na, f.reshape(-1))
    th = 2 * np.pi * (np.arange(nr * na) % na) / na
    assert np.allclose(f[0], np.cos(th), atol=1e-15)
    assert np.allclose(f[1], np.sin(th), atol=1e-15)


def test_weight_bilinear_exact(orc):
    """CIC interpolation reproduces a field bilinear in (u, t) exactly."""
    nr, na = 20, 32
    a = orc.alpha(na)
    rng = np.random.default_rng(5)
    n = 2000
    r = np.exp(a * rng.uniform(0.5, nr - 1.5, n)) - 1
    ph = rng.uniform(0, 2 * np.pi, n)
    x, y = r * np.cos(ph), r * np.sin(ph)
    cell, w, rr = orc.weight(x, y, nr, na)
    u = np.log1p(rr) / a
    t = np.mod(np.arctan2(y, x) * na / (2 * np.pi), na)
    ju, jt = cell // na, cell % na
    assert np.all(ju == np.floor(u).astype(int))
    assert np.all(jt == np.floor(t).astype(int))
    assert np.allclose(w.sum(1), 1, atol=1e-14)
    # f(u, t) = 3 + 2u - 0.5 t + 0.25 u t within the cell's own frame
    du, dt = u - ju, t - jt
    f = lambda U, T: 3 + 2 * U - 0.5 * T + 0.25 * U * T
    interp = (w[:, 0] * f(0, 0) + w[:, 1] * f(0, 1) + w[:, 2] * f(1, 0) + w[:, 3] * f(1, 1))
    assert np.allclose(interp, f(du, dt), atol=1e-12)


def test_assignment_interpolation_transpose(orc):
    """<massgn(p), f> == <p, grdacc(f)>: the same weights and corners feed
    the scatter (massgn.f:68-89) and the gather (grdacc.f:40-66)."""
    nr, na = 16, 32
    a = orc.alpha(na)
    rng = np.random.default_rng(9)
    n = 3000
    r = np.exp(a * rng.uniform(0.2, nr - 1.2, n)) - 1
    ph = rng.uniform(-np.pi, np.pi, n)
    x, y = r * np.cos(ph), r * np.sin(ph)
    iz = np.ones(n, np.int32)
    mesh = nr * na
    g = np.zeros(mesh)
    mfx = np.zeros(mesh, np.int64)
    jr, kr = np.array([nr], np.int32), np.array([0], np.int32)
    orc._lib.orc_massgn_d(n, x, y, iz, nr, na, a, 1, 1, g, mfx, 40, jr, kr)
    f = rng.standard_normal(mesh)
    cell, w, _ = orc.weight(x, y, nr, na)
    c1 = cell
    c2 = np.where((cell + 1) % na == 0, cell + 1 - na, cell + 1)
    gath = w[:, 0] * f[c1] + w[:, 1] * f[c2] + w[:, 2] * f[c1 + na] + w[:, 3] * f[c2 + na]
    assert abs(np.dot(g, f) - gath.sum()) < 1e-10 * abs(gath).sum()
    assert abs(mfx.sum() * 2.0**-40 - n) < 1e-6
    assert jr[0] == (cell // na).min() + 1 and kr[0] == (cell // na).max() + 2


def test_fixforce_hermite_exact_on_cubic(orc):
    dr = 0.25
    rk = dr * np.arange(40)
    val = 1 - 2 * rk + 0.3 * rk**2 - 0.01 * rk**3
    der = -2 + 0.6 * rk - 0.03 * rk**2
    r = np.linspace(0, rk[-1] - 1e-9, 777)
    got = orc.fixforce(r, dr, val, der)
    assert np.allclose(got, 1 - 2 * r + 0.3 * r**2 - 0.01 * r**3, atol=1e-12)


def _harmonic_model(orc, n, omega2, prec="d"):
    nr, na = 64, 64
    m = orc.P2D(nr, na, 4, 1.0, prec=prec, pmass=0.0,
                fixforce=None)
    # fixed radial force F(r) = -omega^2 r (exact for cubic Hermite)
    dr = m.rgrid / 100
    rk = dr * np.arange(120)
    m.fixforce = (dr, -omega2 * rk, -omega2 * np.ones_like(rk))
    return m


def test_leapfrog_harmonic_closed_form(orc):
    """RVLF leap-frog (stpgrp.f:27-33) on F = -w^2 r with a zero grid field:
    x_{n+1} - 2 cos(theta) x_n + x_{n-1} = 0, 2 cos(theta) = 2 - w^2."""
    w2 = 0.01
    m = _harmonic_model(orc, 1, w2)
    x0, v0 = 5.0, 0.3                     # v0 = velocity lagging half a step
    m.load([x0], [0.0], [v0], [0.0])
    m.masset()
    xs = [x0]
    zero = np.zeros((2, m.mesh))
    for _ in range(200):
        m.step(field=zero)
        xs.append(float(m.x[0]))
    th = np.arccos(1 - w2 / 2)
    # x_1 = x0 + v0 - w2 x0; closed form x_n = A cos(n th) + B sin(n th)
    A = x0
    x1 = x0 + v0 - w2 * x0
    B = (x1 - A * np.cos(th)) / np.sin(th)
    ref = A * np.cos(th * np.arange(201)) + B * np.sin(th * np.arange(201))
    assert np.abs(np.array(xs) - ref).max() < 1e-10


def test_chstep_identity_and_zone_scaling(orc):
    """Two zones, particle kept in zone 2: accelerations are scaled by
    nstep^2 and moves happen every second step only."""
    m = _harmonic_model(orc, 1, 0.0)
    m.nzones, m.nstep, m.zbr = 2, [1, 2], [0.0, 1.0]
    m.jrad = np.full(2, m.nr, np.int32)
    m.krad = np.zeros(2, np.int32)
    m.lstep = np.zeros((2, 2), np.int64)
    m.mstep = np.zeros(2, np.int64)
    m.grdmss = np.zeros((2, m.mesh))
    m.grdmoz = np.zeros((2, m.mesh))
    m.mfx = np.zeros((2, m.mesh), np.int64)
    m.mozfx = np.zeros((2, m.mesh), np.int64)
    m.load([10.0], [0.0], [0.0], [0.2], iz=[2])
    m.masset()
    zero = np.zeros((2, m.mesh))
    m.step(field=zero)               # istep 0: zone 2 moves
    assert m.iz[0] == 2 and np.isclose(m.y[0], 0.2)
    y1 = float(m.y[0])
    m.step(field=zero)               # istep 1: zone 2 frozen
    assert float(m.y[0]) == y1
    m.step(field=zero)
    assert np.isclose(m.y[0], 0.4)


def test_fixed_potentials_match_their_forces():
    """The diagnostics' fixed potentials: their derivative is minus the
    fixed force the step applies (central differences on the tables'
    grids), for the exponential-disc workload's halo and the softened
    Kuz'min-Toomre disc; the Freeman disc potential is -md at the centre."""
    from galaxy_amd import disc
    for pot, force in (
            (lambda: disc.halo_potential_table(20.0, 3.3, 0.025, n=2049),
             lambda: disc.halo_table(20.0, 3.3, 0.025, n=2049)),
            (lambda: disc.soft_kt_potential_table(90.0, 10.0, 0.02, 1.367879, n=2049),
             lambda: disc.soft_kt_table(90.0, 10.0, 0.02, 1.367879, n=2049))):
        dr, val, der = pot()
        dr2, fval, _ = force()
        assert dr == dr2
        num = (val[2:] - val[:-2]) / (2 * dr)
        scale = np.abs(fval).max()
        # central differences: O(dr^2) error, largest inside the halo core
        assert np.abs(num - der[1:-1]).max() < 1e-3 * scale
        assert np.abs(der + fval).max() <= 1e-15 * scale
    assert abs(disc.disc_potential(np.array([0.0]))[0] + 0.5) < 1e-15
    R = np.linspace(0.05, 8, 400)
    h = 1e-5
    dphi = (disc.disc_potential(R + h) - disc.disc_potential(R - h)) / (2 * h)
    assert np.abs(dphi + disc.disc_force(R)).max() < 1e-8


def test_oracle_integrals_closed_form(orc):
    """Zone scaling and the kinematic sums of the integrals on hand-placed
    particles: a zone-2 particle stores twice its per-step velocity."""
    m = orc.P2D(16, 32, 2, 0.5, nzones=2, nstep=(1, 2), zbr=(0.0, 3.0), pmass=0.25,
                prec="d")
    x = np.array([1.0, 5.0, -2.0])
    y = np.array([0.0, 0.0, 1.0])
    vx = np.array([0.0, 0.0, 0.3])
    vy = np.array([0.5, 2.0, 0.0])             # particle 2 is in zone 2
    m.load(x, y, vx, vy, np.array([1, 2, 1], np.int32))
    out = m.integrals()
    vy1 = np.array([0.5, 1.0, 0.0])
    assert np.isclose(out[0], 0.75)
    assert np.allclose(out[1:3], 0.25 * np.array([4.0, 1.0]))
    assert np.allclose(out[3:5], 0.25 * np.array([0.3, 1.5]))
    assert np.isclose(out[5], 0.25 * (0.25 + 1.0 + 0.09))
    assert np.isclose(out[6], 0.25 * (1.0 * 0.5 + 5.0 * 1.0 - 1.0 * 0.3))
