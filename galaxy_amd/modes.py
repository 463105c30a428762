"""Density-transform analysis and the fit of one growing mode.

danl() forms, for every ring of a p2d grid, the azimuthal Fourier
coefficients of the combined mass relative to its m = 0 term, as the
reference's DANL analysis does (SRC/lib15/densty.f:217-266).  That
analysis transforms the ring with p2manl and normalises by the m = 0 term.
The ring area and the surface-density factor applied before it
(densty.f:100-110) cancel in that ratio, so they are left out here.

fit_mode() fits c(t, R) = a(R) exp(lambda t) to such a time series.  This
is the growing-mode fit that the reference's modefit program performs on
its .res data (SRC/progs/modefit.f:1-25).  The complex eigenfrequency is
reported as omega = m Omega_p + i s, with pattern speed Omega_p and growth
rate s, the form of the linear-theory value 0.540 + 0.158i that
EXAMPLES/p2d/README quotes for the cold Kuz'min-Toomre disc.  The fit is a
variable projection: for a trial lambda the shape a(R) follows by linear
least squares, and lambda minimises what is left.  Everything here runs on
the host on arrays the library hands back; none of it is on the step's hot
path.
"""
from __future__ import annotations

import numpy as np


def danl(rho, nr: int, na: int, ng: int, nsect: int = 1, skip=()) -> np.ndarray:
    """Relative Fourier coefficients c[ring, m], m = 0..ng-1, of the ring
    masses rho[nr * na] (ring-major): c = sum_j x_j e^{i m theta_j} / sum_j x_j,
    theta_j = 2 pi j / na.  FFTPACK's forward transform gives
    r(2m) = sum x cos, r(2m+1) = -sum x sin, and densty.f stores
    (r(2m), -r(2m+1)) / c0.  Harmonics that the grid's filter removes
    (m not a multiple of nsect, or listed in skip, except m = 0, which the
    analysis turns back on, densty.f:224-228) are zero, as is a ring with
    no mass or whose m = 0 term is below 1e-6 of the harmonic's largest part.
    """
    x = np.asarray(rho, np.float64).reshape(nr, na)
    ms = np.arange(ng)
    th = 2.0 * np.pi * np.outer(np.arange(na), ms) / na      # [na, ng]
    cs = x @ np.cos(th)                                       # r(2m)
    sn = x @ np.sin(th)                                       # -r(2m+1)
    c0 = cs[:, 0]
    out = np.zeros((nr, ng), np.complex128)
    out[:, 0] = np.where(c0 > 0, 1.0, 0.0)
    for m in range(1, ng):
        if m % nsect != 0 or m in skip:
            continue
        big = np.maximum(np.abs(cs[:, m]), np.abs(sn[:, m]))
        ok = (c0 > 0) & (c0 > 1e-6 * big)
        out[ok, m] = (cs[ok, m] + 1j * sn[ok, m]) / c0[ok]
    return out


def _project(lam, t, c):
    """Best shape a[R] for exp(lam t) and the residual sum of squares."""
    e = np.exp(lam * t)                                       # [nt]
    den = np.vdot(e, e).real
    a = (np.conj(e) @ c) / den                                # [nR]
    r = c - np.outer(e, a)
    return a, float(np.vdot(r, r).real)


def fit_mode(t, c, guess=None):
    """Fit c[k, R] ~ a[R] exp(lam t[k]) with lam = s + i m Omega_p.

    Returns (omega, a, rel_resid): omega = m Omega_p + i s (= -i lam
    conjugated into the reference's sign convention, i.e. Im(lam) + i
    Re(lam)), the shape a over the columns of c, and the residual relative
    to |c|^2.  The start value comes from the amplitude-weighted slopes of
    log|c| and of the unwrapped phase unless guess (an omega) is given.
    """
    from scipy.optimize import minimize
    t = np.asarray(t, np.float64)
    c = np.asarray(c, np.complex128)
    if c.ndim == 1:
        c = c[:, None]
    if guess is None:
        w = np.abs(c[-1])
        amp = np.log(np.maximum(np.abs(c), 1e-300))
        ph = np.unwrap(np.angle(c), axis=0)
        A = np.vstack([t, np.ones_like(t)]).T
        s_r = np.linalg.lstsq(A, amp, rcond=None)[0][0]
        f_r = np.linalg.lstsq(A, ph, rcond=None)[0][0]
        lam0 = complex(np.sum(w * s_r) / np.sum(w), np.sum(w * f_r) / np.sum(w))
    else:
        lam0 = complex(guess.imag, guess.real)

    def obj(p):
        return _project(complex(p[0], p[1]), t, c)[1]

    res = minimize(obj, [lam0.real, lam0.imag], method="Nelder-Mead",
                   options=dict(xatol=1e-7, fatol=1e-14 * max(obj([lam0.real, lam0.imag]), 1e-300),
                                maxiter=4000))
    lam = complex(res.x[0], res.x[1])
    a, rss = _project(lam, t, c)
    tot = float(np.vdot(c, c).real)
    return complex(lam.imag, lam.real), a, rss / tot if tot > 0 else 0.0
