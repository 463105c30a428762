"""danl and fit_mode (galaxy_amd/modes.py) against closed forms."""
import numpy as np

from galaxy_amd import modes


def test_danl_closed_form():
    # ring masses 1 + A cos(2 (theta - p0)): c_2 = (A / 2) e^{2 i p0}, the
    # filtered harmonic 1 and the empty ring are zero
    nr, na, ng = 3, 64, 3
    th = 2 * np.pi * np.arange(na) / na
    A, p0 = 0.3, 0.7
    rho = np.zeros((nr, na))
    rho[0] = 1 + A * np.cos(2 * (th - p0))
    rho[1] = 2 * (1 + 0.1 * np.cos(th)) * (1 + A * np.cos(2 * (th - p0)))
    c = modes.danl(rho.ravel(), nr, na, ng, nsect=2)
    want = A / 2 * np.exp(2j * p0)
    assert abs(c[0, 2] - want) < 1e-14
    assert abs(c[1, 2] - want) < 1e-14      # m = 1 part does not leak into m = 2
    assert c[0, 1] == 0 and c[1, 1] == 0
    assert np.all(c[2] == 0)
    assert c[0, 0] == 1 and c[1, 0] == 1


def test_danl_skip_and_sector():
    nr, na, ng = 1, 32, 5
    th = 2 * np.pi * np.arange(na) / na
    rho = 1 + 0.2 * np.cos(4 * th) + 0.1 * np.sin(2 * th)
    c = modes.danl(rho, nr, na, ng, nsect=1, skip=(2,))
    assert c[0, 2] == 0                                  # skipped
    assert abs(c[0, 4] - 0.1) < 1e-14                    # (0.2 / 2) e^0
    assert abs(c[0, 1]) < 1e-15 and abs(c[0, 3]) < 1e-15


def test_fit_mode_recovers_growing_mode():
    rng = np.random.default_rng(3)
    lam = complex(0.158, 0.540)          # growth 0.158, m Omega_p 0.540
    t = np.arange(21.0, 56.0, 0.4)
    a = (0.5 + rng.random(20)) * np.exp(1j * rng.random(20) * 6)
    c = np.outer(np.exp(lam * t), a) * 1e-4
    c = c + 1e-7 * (rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape))
    om, shape, res = modes.fit_mode(t, c)
    assert abs(om - complex(0.540, 0.158)) < 1e-4
    assert res < 1e-6
    # the shape comes back up to the normalisation of the time origin
    assert np.allclose(shape / shape[0], a / a[0], rtol=1e-3)


def test_fit_mode_decaying_and_retrograde():
    t = np.linspace(0, 10, 60)
    lam = complex(-0.05, -1.2)
    c = np.exp(lam * t)[:, None] * np.array([1.0, 2.0j])
    om, _, res = modes.fit_mode(t, c)
    assert abs(om - complex(-1.2, -0.05)) < 1e-6 and res < 1e-12
