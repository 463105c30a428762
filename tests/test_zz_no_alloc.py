"""galhip_integrals_noalloc (static device scratch, no allocation) against
galhip_integrals on the same contexts: bit-identical results, two calls in
a row with the second on a smaller particle set (the scratch then still
holds the first call's larger partial-sum table), on the 2-D grid in FP32
and FP64 and on the 3-D grid."""
import numpy as np
import pytest

from tests._model import SMALL, gpu_model

pytestmark = pytest.mark.gpu


def _vec(d):
    keys = [k for k in d if k != "cm"]
    return np.concatenate([np.atleast_1d(np.asarray(d[k], np.float64)) for k in sorted(keys)])


def _same(a, b):
    va, vb = _vec(a), _vec(b)
    assert np.array_equal(va, vb, equal_nan=True), (va, vb)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_p2d_noalloc_matches_integrals(prec):
    from galaxy_amd import disc
    print("noalloc p2d", prec, flush=True)
    for n in (40_000, 12_345):
        g = gpu_model(precision=prec, n=n)
        fp = disc.halo_potential_table(float(g.rgrid), SMALL.lscale(), SMALL.dt)
        g.set_fixpot(*fp)
        g.masset()
        g.findf(False)
        _same(g.integrals(centred=False), g.integrals(centred=False, noalloc=True))
        assert np.isnan(g.integrals(centred=False, noalloc=True)["w_grid"])
        g.run(3)
        g.findf(True)
        ref = g.integrals(centred=True)
        got = g.integrals(centred=True, noalloc=True)
        _same(ref, got)
        assert np.isfinite(got["w_grid"]) and got["mass"] > 0
        g.close()


def test_p3d_noalloc_matches_integrals():
    from galaxy_amd import configs
    from galaxy_amd.lib15 import P3DGrid
    print("noalloc p3d", flush=True)
    for n in (30_000, 7_000):
        w = configs.Workload3("p3d-24x64x16-noalloc", nr=24, na=64, hash=4, n=n, steps=1,
                              dt=0.05, ngz=16, dz_model=0.1, eps_model=0.2, seed=5)
        kw = configs.grid_kwargs3(w)
        g = P3DGrid(precision="fp32", **kw)
        g.greenm()
        g.restrt(*configs.particles3(w, None, None, g.rgrid))
        g.masset()
        g.findf(True)
        ref = g.integrals(centred=True)
        got = g.integrals(centred=True, noalloc=True)
        _same(ref, got)
        assert np.isfinite(got["w_grid"]) and got["mass"] > 0
        g.close()
