"""Builds matching oracle and GPU models of a small three-zone disc, for the
parity tests.  Grid 64 x 128 (outer cell / softening = 2.9, inside the
reference's sanity bound, SRC/lib15/setgrd.f:182-198), m <= 8, zones with
nstep 1, 2, 4."""
import numpy as np

from galaxy_amd import configs

SMALL = configs.Workload("p2d-64x128-test", nr=64, na=128, hash=8, n=40_000,
                         steps=8, dt=0.05, nstep=(1, 2, 4),
                         zbr_model=(0.0, 1.5, 3.5), seed=11)


def oracle_model(orc, w=SMALL, prec="d", n=None, seed=None, G=None):
    kw = configs.grid_kwargs(w, n)
    m = orc.P2D(kw["nr"], kw["na"], kw["hash_"], kw["softl"], tsoft=1,
                nsect=kw["nsect"], nzones=kw["nzones"], nstep=kw["nstep"],
                zbr=kw["zbr"], pmass=kw["pmass"], offanal=True, prec=prec,
                use_fx=True, green_tables=G)
    m.fixforce = configs.fixforce(w, float(m.rgrid))
    x, y, vx, vy, iz = configs.particles(w, n, seed, float(m.rgrid))
    m.load(x, y, vx, vy, iz)
    return m, (x, y, vx, vy, iz)


def gpu_model(w=SMALL, precision="fp32", n=None, seed=None, parts=None):
    from galaxy_amd.lib15 import P2DGrid
    g = P2DGrid(precision=precision, **configs.grid_kwargs(w, n))
    g.greenm()
    dr, val, der = configs.fixforce(w, g.rgrid)
    g.set_fixforce(dr, val, der)
    if parts is None:
        parts = configs.particles(w, n, seed, g.rgrid)
    g.restrt(*parts)
    return g


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
