"""Host parts of the gradient pipeline, on the CPU: the |A_2|^2 loss and its
position gradient, and the initial-velocity derivative in Q and md, each
against central differences and the defining formulas."""
import numpy as np

from galaxy_amd import adjoint, configs, disc


def test_a2_loss_value_and_gradient():
    rng = np.random.default_rng(1)
    x = rng.normal(size=3000) * 8
    y = rng.normal(size=3000) * 8
    ls = 3.6
    L, gx, gy = adjoint.a2_loss(x, y, ls)
    R = np.hypot(x, y) / ls
    sel = (R >= 1) & (R <= 3)
    A = np.exp(-2j * np.arctan2(y[sel], x[sel])).sum() / sel.sum()
    assert abs(L - abs(A) ** 2) <= 1e-14
    d, e, h = rng.normal(size=3000), rng.normal(size=3000), 1e-6
    fd = (adjoint.a2_loss(x + h * d, y + h * e, ls, sel=sel)[0]
          - adjoint.a2_loss(x - h * d, y - h * e, ls, sel=sel)[0]) / (2 * h)
    assert abs(fd - (gx @ d + gy @ e)) <= 1e-6 * abs(fd)
    # particles outside the annulus carry no gradient
    assert np.all(gx[~sel] == 0) and np.all(gy[~sel] == 0)


def test_ic_velocities_reproduce_the_generator_and_their_q_derivative():
    w = configs.Workload("t", nr=64, na=128, hash=8, n=5000, steps=1, dt=0.05, Q=1.3, seed=9)
    R, phi, g1, g2 = disc.disc_draws(w.n, w.seed, w.rtrunc)
    terms = disc.velocity_terms(R, w.rtrunc, w.rcore)
    cp, sp, f = np.cos(phi), np.sin(phi), 2.0
    vx, vy = adjoint.ic_velocities(*terms, g1, g2, cp, sp, w.Q, w.md, f)
    _, _, vR, vphi = disc.exponential_disc(w.n, w.seed, Q=w.Q, md=w.md)
    assert np.abs(vx - (vR * cp - vphi * sp) * f).max() < 1e-13
    assert np.abs(vy - (vR * sp + vphi * cp) * f).max() < 1e-13
    dx, dy = adjoint.ic_velocity_dq(*terms, g1, g2, cp, sp, w.Q, w.md, f)
    h = 1e-6
    p = adjoint.ic_velocities(*terms, g1, g2, cp, sp, w.Q + h, w.md, f)
    m = adjoint.ic_velocities(*terms, g1, g2, cp, sp, w.Q - h, w.md, f)
    assert np.abs((p[0] - m[0]) / (2 * h) - dx).max() < 1e-7
    assert np.abs((p[1] - m[1]) / (2 * h) - dy).max() < 1e-7
    # md enters through the same product Q md
    pm = adjoint.ic_velocities(*terms, g1, g2, cp, sp, w.Q, w.md + h, f)
    mm = adjoint.ic_velocities(*terms, g1, g2, cp, sp, w.Q, w.md - h, f)
    assert np.abs((pm[0] - mm[0]) / (2 * h) - dx * w.Q / w.md).max() < 1e-7
