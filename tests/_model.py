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


# ---- the reference's standard p2d test: EXAMPLES/p2d/run100 ----------------
#
# Cold, minimally softened Kuz'min-Toomre disc (disc type SOFT, epsilon =
# 0.367879), 15000 particles placed quietly, 93 x 128 grid, harmonic m = 2
# only (HASH 2, sectors 2, SKIP 0 replaces m = 0 by the fixed axisymmetric
# force), two zones (step ratio 2 beyond 3 scale lengths = 30 grid units),
# dt = 0.02, lscale 10 (EXAMPLES/p2d/run100.dat).  Particles and scales
# are those of the reference's run100.dmp0 (tests/golden/run100_dmp0.npz).
# Its dominant mode has eigenfrequency 0.540 + 0.158i by linear theory, and
# 0.5392 + 0.1584i in the reference author's run (EXAMPLES/p2d/README).
KT_OMEGA = complex(0.540, 0.158)


def kt_setup():
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "run100_dmp0.npz"))
    rgrid, lscale, ts, pmass, softl = (float(v) for v in d["reals"])
    kw = dict(nr=93, na=128, hash_=2, softl=softl, tsoft=1, nsect=2, skip=(0,),
              nzones=2, nstep=(1, 2), zbr=(0.0, 3.0 * lscale), offanal=True,
              pmass=pmass)
    from galaxy_amd import disc
    ff = disc.soft_kt_table(rgrid, lscale, ts, 1.0 + 0.367879)
    parts = (d["x"], d["y"], d["vx"], d["vy"], d["iz"].astype(np.int32))
    return kw, ff, parts, dict(rgrid=rgrid, lscale=lscale, ts=ts)


def kt_mode(advance, nsteps=2800, every=20, window=(21.0, 56.0)):
    """Runs advance(analyse) nsteps times (one findf + step each; it returns
    the combined mass grid when analyse is set, every `every` steps as the
    run's analysis interval), forms the m = 2 density transform (DANL) and
    fits one growing mode over the time window to the rings between 0.2 and
    3 scale lengths.  Returns (omega, relative residual, amplitude at the
    window's end)."""
    from galaxy_amd import modes
    kw, _, _, sc = kt_setup()
    nr, na = kw["nr"], kw["na"]
    T, C = [], []
    for s in range(nsteps):
        an = s % every == 0
        rho = advance(an)
        if an:
            T.append(s * sc["ts"])
            C.append(modes.danl(rho, nr, na, 3, nsect=2, skip=(0,))[:, 2])
    T, C = np.array(T), np.array(C)
    R = (np.exp(2.0 * np.pi / na * np.arange(nr)) - 1.0) / sc["lscale"]
    sel = (R > 0.2) & (R < 3.0)
    w = (T >= window[0]) & (T <= window[1])
    om, _, res = modes.fit_mode(T[w], C[w][:, sel])
    return om, res, float(np.abs(C[w][-1, sel]).max())
