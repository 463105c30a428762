"""Compensated position updates (galhip_set_compensated).

No counterpart in the reference, which updates FP32 positions plainly; the
checks are what the mathematics fixes:
  - free drift (no force): after K steps x0 + K v in exact arithmetic; the
    compensated hi + lo stays within an ulp of it while the plain update
    accumulates rounding (the plain run is the reference's arithmetic);
  - graph replay and re-sorts carry the low words (bit-identical runs);
  - an ensemble member with compensation equals its lone disc with it.
"""
import numpy as np
import pytest

from galaxy_amd import configs
from galaxy_amd.lib15 import P2DGrid
from tests._model import SMALL, gpu_model

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _drift_grid(comp, n=4000, seed=3):
    kw = configs.grid_kwargs(SMALL, n)
    kw.update(nzones=1, nstep=(1,), zbr=(0.0,))
    g = P2DGrid(precision="fp32", **kw)
    g.greenm()
    rng = np.random.default_rng(seed)
    r = rng.uniform(2.0, 0.5 * g.rgrid, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    x = (r * np.cos(ph)).astype(np.float32)
    y = (r * np.sin(ph)).astype(np.float32)
    vx = rng.uniform(-3e-3, 3e-3, n).astype(np.float32)
    vy = rng.uniform(-3e-3, 3e-3, n).astype(np.float32)
    g.set_compensated(comp)
    g.restrt(x, y, vx, vy, np.ones(n, np.int32))
    g.set_field(np.zeros((3, g.mesh), np.float32))
    return g, x, y, vx, vy


def test_free_drift():
    K = 1500
    _say("free drift", K)
    errs = {}
    for comp in (False, True):
        g, x, y, vx, vy = _drift_grid(comp)
        for _ in range(K):
            g.step()
        out = g.unload()
        xl, yl = g.get_position_lo()
        ex = x.astype(np.float64) + K * vx.astype(np.float64)
        ey = y.astype(np.float64) + K * vy.astype(np.float64)
        px = out["x"].astype(np.float64) + xl.astype(np.float64)
        py = out["y"].astype(np.float64) + yl.astype(np.float64)
        ulp = np.spacing(np.maximum(np.abs(out["x"]), np.abs(out["y"])).astype(np.float32))
        err = np.maximum(np.abs(px - ex), np.abs(py - ey)) / ulp
        errs[comp] = float(err.max())
        # the velocities never change without a force
        assert np.array_equal(out["vx"], vx) and np.array_equal(out["vy"], vy)
        if not comp:
            assert np.all(xl == 0) and np.all(yl == 0)
        g.close()
    _say("max error in ulps: plain %.3g, compensated %.3g" % (errs[False], errs[True]))
    assert errs[True] <= 1.0
    assert errs[False] > 10 * max(errs[True], 0.01)


def test_graph_and_sort_carry_low_words():
    _say("compensated graph replay")
    A = gpu_model(precision="fp32", n=30000)
    B = gpu_model(precision="fp32", n=30000)
    for g in (A, B):
        g.set_compensated(True)
        g.restrt(*configs.particles(SMALL, 30000, None, g.rgrid))
        g.set_sort_interval(3)
        g.masset()
        g.findf()
    A.set_graph(True)
    for k in range(2):
        A.run(8)
        B.run(8)
        ua, ub = A.unload(), B.unload()
        for key in ("x", "y", "vx", "vy", "iz"):
            assert np.array_equal(ua[key], ub[key]), (k, key)
        la, lb = A.get_position_lo(), B.get_position_lo()
        assert np.array_equal(la[0], lb[0]) and np.array_equal(la[1], lb[1]), k
        assert np.any(la[0] != 0)
    A.close()
    B.close()


def test_ensemble_member_equals_lone_disc_compensated():
    from tests.test_gpu_ensemble import COUNTS, _ensemble, _singles
    _say("compensated ensemble")
    E, parts = _ensemble("fp32")
    E.set_compensated(True)
    E.restrt(parts)
    S = _singles(E, parts, "fp32")
    for s in S:
        s.set_compensated(True)
    for i, s in enumerate(S):
        s.restrt(*parts[i])
    for g in [E] + S:
        g.masset()
        g.findf()
        g.run(7)
    m = E.unload()
    for i, s in enumerate(S):
        o = s.unload()
        for key in ("x", "y", "vx", "vy", "iz"):
            assert np.array_equal(m[i][key], o[key]), (i, key)
    lo = E.get_position_lo()
    off = np.concatenate([[0], np.cumsum(COUNTS)])
    for i, s in enumerate(S):
        sl = s.get_position_lo()
        assert np.array_equal(lo[0][off[i]:off[i + 1]], sl[0]), i
    for g in [E] + S:
        g.close()
