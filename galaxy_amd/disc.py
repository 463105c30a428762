"""Seeded exponential-disc initial conditions and the rigid-halo force table.

This replaces, for the workloads measured here, the reference's set-up
programs (mkpcs/psetup/qsetup/orbset, SRC/progs/mkpcs.f:98-118,
SRC/lib15/qsetup.f:4-22) with a seeded generator, so identical particles
can be made on any machine.

Model (units G = 1, disc scale length R_d = 1):
  surface density  Sigma(R) = Sigma0 exp(-R), truncated at R = rtrunc,
                   disc mass md (Sigma0 fixed by md and the truncation)
  rotation curve   v_c^2(R) = R^2 / (R^2 + rcore^2)  (flat at 1 outside a
                   small core that keeps the centre regular)
  rigid halo       F_halo = -v_c^2/R - F_disc(R), the reference's
                   fixed-rotation-curve rule (SRC/lib15/frtot.f:46 minus
                   SRC/lib15/frdisc.f:50, the untruncated exponential disc:
                   F_disc = -(md/R_d^2) y (I0K0 - I1K1), y = R / 2)
  velocities       sigma_R = 3.36 Q Sigma / kappa, sigma_phi^2 =
                   sigma_R^2 kappa^2 / (4 Omega^2), mean v_phi from the
                   Jeans asymmetric-drift relation (the recipe qsetup
                   describes), Gaussian deviates
Positions: R from the inverse enclosed-mass CDF (Newton in FP64), phi
uniform.  Random numbers: numpy Philox keyed by the seed.

Internal units (SRC/lib15/setrun.f:858-873): lengths x lscale (grid units),
one step = ts model time units; a velocity is a displacement per step and
is further multiplied by nstep(zone) for particles in slower zones
(SRC/lib15/adjvls.f); pmass = md lscale^3 ts^2 / N.
"""
from __future__ import annotations

import math

import numpy as np
from scipy import special


def _bessel_terms(y):
    i0, i1 = special.i0e(y), special.i1e(y)
    k0, k1 = special.k0e(y), special.k1e(y)
    # I_n(y) K_m(y) = I_n e^-y * K_m e^y
    return i0 * k0, i1 * k1, i1 * k0, i0 * k1


def disc_force(R, md=0.5):
    """F_disc(R) of the untruncated exponential disc (frdisc 'EXP ')."""
    R = np.asarray(R, np.float64)
    y = 0.5 * np.abs(R)
    with np.errstate(invalid="ignore", divide="ignore"):
        a, b, _, _ = _bessel_terms(np.where(y > 0, y, 1.0))
        f = -md * y * (a - b)
    return np.where(y > 0, f, 0.0)


def disc_force_deriv(R, md=0.5):
    """dF_disc/dR: -(md/2) [I0K0 + I1K1 + 2y (I1K0 - I0K1)]."""
    R = np.asarray(R, np.float64)
    y = np.maximum(0.5 * np.abs(R), 1e-300)
    a, b, c, d = _bessel_terms(y)
    return -0.5 * md * (a + b + 2.0 * y * (c - d))


def halo_force(R, md=0.5, rcore=0.1):
    """Radial force of the rigid halo in model units."""
    R = np.asarray(R, np.float64)
    return -R / (R * R + rcore * rcore) - disc_force(R, md)


def halo_force_deriv(R, md=0.5, rcore=0.1):
    R = np.asarray(R, np.float64)
    s = R * R + rcore * rcore
    return -(rcore * rcore - R * R) / (s * s) - disc_force_deriv(R, md)


def halo_table(rgrid, lscale, ts, md=0.5, rcore=0.1, n=4097):
    """Fixed radial force table in internal units on r_k = k dr covering
    [0, rgrid] (grid units): values and derivatives for cubic Hermite
    interpolation."""
    dr = rgrid / (n - 2)
    r = dr * np.arange(n)
    R = r / lscale
    val = halo_force(R, md, rcore) * lscale * ts * ts
    der = halo_force_deriv(R, md, rcore) * ts * ts
    return dr, val, der


def disc_potential(R, md=0.5):
    """Potential of the untruncated exponential disc (Freeman):
    -(md/2) R [I0(R/2) K1(R/2) - I1(R/2) K0(R/2)], -md at R = 0."""
    R = np.asarray(R, np.float64)
    y = 0.5 * np.abs(R)
    yy = np.where(y > 0, y, 1.0)
    _, _, i1k0, i0k1 = _bessel_terms(yy)
    return np.where(y > 0, -0.5 * md * np.abs(R) * (i0k1 - i1k0), -md)


def halo_potential(R, md=0.5, rcore=0.1):
    """Potential of the rigid halo of halo_force: (1/2) ln(R^2 + rcore^2)
    (the flat-curve total) minus the disc's."""
    R = np.asarray(R, np.float64)
    return 0.5 * np.log(R * R + rcore * rcore) - disc_potential(R, md)


def halo_potential_table(rgrid, lscale, ts, md=0.5, rcore=0.1, n=4097):
    """The halo potential in internal units (velocity^2: (lscale ts)^2 times
    the model value) on the halo_table grid, derivative = -force."""
    dr = rgrid / (n - 2)
    r = dr * np.arange(n)
    R = r / lscale
    val = halo_potential(R, md, rcore) * (lscale * ts) ** 2
    der = -halo_force(R, md, rcore) * lscale * ts * ts
    return dr, val, der


def soft_kt_table(rgrid, lscale, ts, rstar, mass=1.0, n=4097):
    """Fixed radial force of a softened Kuz'min-Toomre disc (disc type SOFT:
    F(R) = -M R / (rstar^2 + R^2)^1.5 with rstar = 1 + epsilon,
    SRC/lib15/frdisc.f:34, inimod.f:74) in internal units, the reference's
    axifrc = frtot(r / lscale) ts / gvfac (SRC/lib15/axifrc.f, gvfac =
    1/(lscale ts), setrun.f:873), tabulated like halo_table."""
    dr = rgrid / (n - 2)
    r = dr * np.arange(n)
    R = r / lscale
    a2 = rstar * rstar
    s = a2 + R * R
    val = -mass * R / s ** 1.5 * lscale * ts * ts
    der = -mass * (a2 - 2.0 * R * R) / s ** 2.5 * ts * ts
    return dr, val, der


def soft_kt_potential_table(rgrid, lscale, ts, rstar, mass=1.0, n=4097):
    """Potential -M / sqrt(rstar^2 + R^2) of soft_kt_table's disc (phidsc
    'SOFT', SRC/lib15/phidsc.f:38) in internal units, derivative = -force."""
    dr = rgrid / (n - 2)
    r = dr * np.arange(n)
    R = r / lscale
    s = rstar * rstar + R * R
    val = -mass / np.sqrt(s) * (lscale * ts) ** 2
    der = mass * R / s ** 1.5 * lscale * ts * ts
    return dr, val, der


def _enclosed_fraction_inverse(u, rtrunc):
    """Solve 1 - (1 + R) e^-R = u (1 - (1 + rt) e^-rt) for R by Newton."""
    mt = 1.0 - (1.0 + rtrunc) * math.exp(-rtrunc)
    target = u * mt
    R = np.where(target < 0.5, np.sqrt(2.0 * target), 2.0 - np.log1p(-target))
    R = np.clip(R, 0.0, rtrunc)
    # quadratic convergence from this start: 1e-9 after 5 steps, then the
    # rounding floor (~1e-13) of f near R = 0
    for _ in range(12):
        f = 1.0 - (1.0 + R) * np.exp(-R) - target
        fp = R * np.exp(-R)
        step = f / np.maximum(fp, 1e-300)
        R = np.clip(R - step, 0.0, rtrunc)
        if np.max(np.abs(step)) < 1e-12:
            break
    return R


def disc_draws(n, seed, rtrunc=5.0):
    """The random part of exponential_disc: radii, angles and the two
    standard normal deviates (vR, vphi) of each particle."""
    rng = np.random.Generator(np.random.Philox(key=seed))
    u = rng.random(n)
    R = _enclosed_fraction_inverse(u, rtrunc)
    phi = rng.random(n) * (2.0 * math.pi)
    g1 = rng.standard_normal(n)
    g2 = rng.standard_normal(n)
    return R, phi, g1, g2


def velocity_terms(R, rtrunc=5.0, rcore=0.1):
    """Per-particle terms of exponential_disc's velocities that do not depend
    on Q or md: sigma_R = Q md k_sig, sigma_phi = sigma_R k_phi,
    mean v_phi^2 = vc2 + sigma_R^2 c (clamped at 0).  Returns
    (k_sig, k_phi, vc2, c)."""
    mt = 1.0 - (1.0 + rtrunc) * math.exp(-rtrunc)
    s2 = R * R + rcore * rcore
    vc2 = R * R / s2
    om2 = 1.0 / s2
    kap2 = 2.0 * (R * R + 2.0 * rcore * rcore) / (s2 * s2)
    k_sig = 3.36 * np.exp(-R) / (2.0 * math.pi * mt) / np.sqrt(kap2)
    k_phi = np.sqrt(kap2 / (4.0 * om2))
    h = 1e-6

    def lnss(Rv):
        s2v = Rv * Rv + rcore * rcore
        k2 = 2.0 * (Rv * Rv + 2.0 * rcore * rcore) / (s2v * s2v)
        return -3.0 * Rv - np.log(k2)
    Rs = np.maximum(R, 1e-8)
    dlog = (lnss(Rs * (1 + h)) - lnss(Rs * (1 - h))) / (2 * h)
    c = dlog + 1.0 - kap2 / (4.0 * om2)
    return k_sig, k_phi, vc2, c


def exponential_disc(n, seed, Q=1.5, md=0.5, rtrunc=5.0, rcore=0.1):
    """Model-unit particles: (R, phi, vR, vphi) arrays, FP64."""
    rng = np.random.Generator(np.random.Philox(key=seed))
    u = rng.random(n)
    R = _enclosed_fraction_inverse(u, rtrunc)
    phi = rng.random(n) * (2.0 * math.pi)
    mt = 1.0 - (1.0 + rtrunc) * math.exp(-rtrunc)
    sigma0 = md / (2.0 * math.pi * mt)
    s2 = R * R + rcore * rcore
    vc2 = R * R / s2
    om2 = 1.0 / s2
    kap2 = 2.0 * (R * R + 2.0 * rcore * rcore) / (s2 * s2)
    kap = np.sqrt(kap2)
    sig = sigma0 * np.exp(-R)
    sR = 3.36 * Q * sig / kap
    sphi2 = sR * sR * kap2 / (4.0 * om2)
    # d ln(Sigma sigma_R^2) / d ln R, with sigma_R ~ Sigma / kappa
    h = 1e-6
    def lnss(Rv):
        s2v = Rv * Rv + rcore * rcore
        k2 = 2.0 * (Rv * Rv + 2.0 * rcore * rcore) / (s2v * s2v)
        return -3.0 * Rv - np.log(k2)
    Rs = np.maximum(R, 1e-8)
    dlog = (lnss(Rs * (1 + h)) - lnss(Rs * (1 - h))) / (2 * h)
    vphi2 = vc2 + sR * sR * (dlog + 1.0 - sphi2 / np.maximum(sR * sR, 1e-300))
    vphi_mean = np.sqrt(np.maximum(vphi2, 0.0))
    vR = sR * rng.standard_normal(n)
    vphi = vphi_mean + np.sqrt(sphi2) * rng.standard_normal(n)
    return R, phi, vR, vphi


def to_internal(R, phi, vR, vphi, lscale, ts, zbr_model, nstep, rgrid):
    """Cartesian internal-unit coordinates and zones (zoneno rule on the
    cylindrical radius, SRC/lib15/zoneno.f:40-55), velocities scaled by
    nstep(zone) (adjvls).  Particles beyond rgrid go to the off-grid list."""
    x = R * np.cos(phi) * lscale
    y = R * np.sin(phi) * lscale
    vx = (vR * np.cos(phi) - vphi * np.sin(phi)) * lscale * ts
    vy = (vR * np.sin(phi) + vphi * np.cos(phi)) * lscale * ts
    r = R * lscale
    nz = len(nstep)
    iz = np.ones(len(R), np.int32)
    for i in range(nz):
        iz[r >= zbr_model[i] * lscale] = i + 1
    fac = np.asarray(nstep, np.float64)[iz - 1]
    vx *= fac
    vy *= fac
    iz[r > rgrid] = nz + 1
    return x, y, vx, vy, iz


def vertical_sech2(R, seed, z0=0.1, md=0.5, rtrunc=5.0):
    """Heights and vertical velocities (model units) for a disc given the
    vertical profile rho ~ sech^2(z / z0) at every radius, with sigma_z from
    the isothermal sheet's equilibrium sigma_z^2 = pi G Sigma(R) z0 (Spitzer
    1942); Gaussian vertical velocities.  Its own Philox stream (key
    seed + 2^32), so (R, phi, vR, vphi) stay those of exponential_disc."""
    rng = np.random.Generator(np.random.Philox(key=seed + (1 << 32)))
    n = len(R)
    u = rng.random(n)
    # inverse of the sech^2 CDF (1 + tanh(z / z0)) / 2
    z = z0 * np.arctanh(np.clip(2.0 * u - 1.0, -1.0 + 1e-15, 1.0 - 1e-15))
    mt = 1.0 - (1.0 + rtrunc) * math.exp(-rtrunc)
    sig = md / (2.0 * math.pi * mt) * np.exp(-R)
    vz = np.sqrt(math.pi * sig * z0) * rng.standard_normal(n)
    return z, vz


def to_internal3(R, phi, z, vR, vphi, vz, lscale, ts, zbr_model, nstep, rgrid, zm):
    """Internal units for the 3-D grid: as to_internal, zones by the
    spherical radius (SRC/lib15/zoneno.f:40-41, setrun.f:810), off the grid
    when R > rgrid or |z| > zm (SRC/lib15/offgrd.f:69-72)."""
    x = R * np.cos(phi) * lscale
    y = R * np.sin(phi) * lscale
    zz = z * lscale
    vx = (vR * np.cos(phi) - vphi * np.sin(phi)) * lscale * ts
    vy = (vR * np.sin(phi) + vphi * np.cos(phi)) * lscale * ts
    vzz = vz * lscale * ts
    r = np.sqrt(R * R + z * z) * lscale
    nz = len(nstep)
    iz = np.ones(len(R), np.int32)
    for i in range(nz):
        iz[r >= zbr_model[i] * lscale] = i + 1
    fac = np.asarray(nstep, np.float64)[iz - 1]
    vx *= fac
    vy *= fac
    vzz *= fac
    iz[(R * lscale > rgrid) | (np.abs(zz) > zm)] = nz + 1
    return x, y, zz, vx, vy, vzz, iz
