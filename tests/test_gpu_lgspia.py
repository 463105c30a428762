"""galhip_lgspia (log-spiral coefficients, FP64 device sums) against the
oracle's restatement on the same particles: FP32 and FP64 p2d grids after
a few steps (the particles as the device holds them, unloaded and passed
to the oracle), odd and even pitch counts, two sectors, offanal on and off;
a p3d grid (the (x, y) of its particles); every member of an ensemble.
Bar: 1e-12 absolute on coefficients of modulus <= 1 (both sum in FP64, in
different orders)."""
import numpy as np
import pytest

from galaxy_amd import configs
from tests._model import SMALL, gpu_model

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _check(orc, out, nz, offanal, nsect, np_, nm, ls, got):
    c, s = orc.lgspia(out["x"], out["y"], out["iz"], nz, offanal, nsect, np_, nm, ls)
    assert np.abs(got[0] - c).max() <= 1e-12
    assert np.abs(got[1] - s).max() <= 1e-12


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_lgspia_p2d(orc, prec):
    _say("lgspia p2d", prec)
    g = gpu_model(precision=prec, n=30000)
    g.masset()
    g.run(5)
    out = g.unload()
    ls = SMALL.lscale()
    for np_, nm in ((61, 4), (60, 3), (1, 8)):
        _check(orc, out, SMALL.nzones, True, 1, np_, nm, ls, g.lgspia(np_, nm, ls))
    g.close()


def test_lgspia_sectors_offgrid(orc):
    """Two sectors, off-grid particles left out (offanal off)."""
    from galaxy_amd.lib15 import P2DGrid
    _say("lgspia sectors, offanal off")
    w = SMALL
    kw = configs.grid_kwargs(w, 20000)
    kw.update(nsect=2, offanal=False)
    g = P2DGrid(precision="fp64", **kw)
    g.greenm()
    x, y, vx, vy, iz = configs.particles(w, 20000, None, g.rgrid)
    x, y = x * 1.3, y * 1.3
    iz = np.where(np.hypot(x, y) > g.rgrid, w.nzones + 1, iz).astype(np.int32)
    g.restrt(x, y, vx, vy, iz)
    out = g.unload()
    assert (out["iz"] == w.nzones + 1).any()
    _check(orc, out, w.nzones, False, 2, 31, 2, w.lscale(), g.lgspia(31, 2, w.lscale()))
    g.close()


def test_lgspia_p3d_and_ensemble(orc):
    from tests.test_gpu_ensemble import _ensemble, COUNTS
    from tests.test_gpu_p3d import P3SMALL
    _say("lgspia p3d and ensemble")
    g = configs.make_grid3(P3SMALL, precision="fp32")
    out = g.unload()
    ls = P3SMALL.lscale()
    _check(orc, out, P3SMALL.nzones, True, 1, 61, 2, ls, g.lgspia(61, 2, ls))
    g.close()
    E, _ = _ensemble("fp64")
    E.masset()
    E.findf()
    E.run(3)
    members = E.unload()
    got = E.lgspia(21, 2, SMALL.lscale())
    for i, m in enumerate(members):
        if COUNTS[i]:
            _check(orc, m, SMALL.nzones, True, 1, 21, 2, SMALL.lscale(), got[i])
        else:
            assert not got[i][0].any() and not got[i][1].any()
    E.close()
