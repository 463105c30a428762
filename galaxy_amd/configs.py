"""Pinned workloads of the polar-grid PM step.

The disc model is the exponential disc of galaxy_amd/disc.py.  Values the
workload descriptions leave open are fixed here and marked "assumed":
  - disc mass md = 0.5 (the active mass fraction), truncation R = 5;
  - rotation-curve core rcore = 0.1 R_d (keeps the halo force finite at
    R = 0; the curve is flat outside it);
  - the grid uses the reference's standard log-polar map (no hole) and
    lscale = rgrid / 6 so that the grid edge is at R = 6 R_d;
  - zone boundaries of the three-zone workload at R = 1.5 and 3.5 R_d
    (assumed);
  - off-grid particles keep moving (offanal true, as in
    EXAMPLES/p2d/run100.dat 'offgrid t');
  - 3-D grid: plane spacing 0.025 R_d (a quarter of the softening, which
    p3dset requires to be at least the spacing), dt = 0.05, one zone
    (assumed; the workload names neither).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import disc


@dataclass
class Workload:
    name: str
    nr: int
    na: int
    hash: int
    n: int
    steps: int
    dt: float                     # model time of one step of zone 1
    nstep: tuple = (1,)
    zbr_model: tuple = (0.0,)     # zone inner radii in units of R_d
    eps_model: float = 0.1        # Plummer softening, units of R_d
    Q: float = 1.5
    seed: int = 1
    md: float = 0.5
    rtrunc: float = 5.0
    rcore: float = 0.1
    redge: float = 6.0

    @property
    def nzones(self):
        return len(self.nstep)

    def rgrid_d(self):
        alpha = 2.0 * math.pi / (2 * (self.na // 2))
        return math.exp(alpha * (self.nr - 1)) - 1.0

    def lscale(self):
        return self.rgrid_d() / self.redge


# Plain method, 128 x 128 grid, N = 1e5, m <= 8, dt = 0.05, 100 steps
C1 = Workload("p2d-128x128-n1e5", nr=128, na=128, hash=8, n=100_000,
              steps=100, dt=0.05, seed=1)
# 256 x 512 grid, N = 1e7, m <= 16, 3 zones dt = 0.025/0.05/0.1, 500 steps
C2 = Workload("p2d-256x512-n1e7-3zones", nr=256, na=512, hash=16,
              n=10_000_000, steps=500, dt=0.025, nstep=(1, 2, 4),
              zbr_model=(0.0, 1.5, 3.5), seed=2)



@dataclass
class Workload3(Workload):
    """A 3-D polar-grid workload: ngz planes dz_model apart (units of R_d),
    disc with a sech^2(z / z0) vertical profile."""
    ngz: int = 125
    dz_model: float = 0.025
    z0: float = 0.1


# 200 x 256 x 125 grid, N = 5e7, m <= 8, sech^2(z/0.1), 200 steps
C3 = Workload3("p3d-200x256x125-n5e7", nr=200, na=256, hash=8, n=50_000_000,
               steps=200, dt=0.05, ngz=125, seed=3)

WORKLOADS = {"C1": C1, "C2": C2, "C3": C3}


def grid_kwargs(w: Workload, n=None):
    """Arguments of P2DGrid / the oracle's P2D for workload w."""
    n = w.n if n is None else n
    ls = w.lscale()
    ts = w.dt
    return dict(nr=w.nr, na=w.na, hash_=w.hash, softl=w.eps_model * ls,
                tsoft=1, nsect=1, nzones=w.nzones, nstep=w.nstep,
                zbr=tuple(z * ls for z in w.zbr_model), offanal=True,
                pmass=w.md * ls ** 3 * ts * ts / n)


def particles(w: Workload, n=None, seed=None, rgrid=None):
    """Internal-unit particles (FP64 arrays) and zones for workload w."""
    n = w.n if n is None else n
    seed = w.seed if seed is None else seed
    R, phi, vR, vphi = disc.exponential_disc(n, seed, Q=w.Q, md=w.md,
                                             rtrunc=w.rtrunc, rcore=w.rcore)
    rg = w.rgrid_d() if rgrid is None else rgrid
    return disc.to_internal(R, phi, vR, vphi, w.lscale(), w.dt, w.zbr_model,
                            w.nstep, rg)


def fixforce(w: Workload, rgrid=None):
    rg = w.rgrid_d() if rgrid is None else rgrid
    return disc.halo_table(rg, w.lscale(), w.dt, md=w.md, rcore=w.rcore)


def fixpot(w: Workload, rgrid=None):
    """The halo's potential (energy diagnostics), on fixforce's grid."""
    rg = w.rgrid_d() if rgrid is None else rgrid
    return disc.halo_potential_table(rg, w.lscale(), w.dt, md=w.md, rcore=w.rcore)


def make_grid(w: Workload, precision="fp32", n=None, seed=None, green=True):
    """A ready P2DGrid for workload w: Green's tables, halo table, particles
    loaded and assigned (masset)."""
    from .lib15 import P2DGrid
    g = P2DGrid(precision=precision, **grid_kwargs(w, n))
    if green:
        g.greenm()
    dr, val, der = fixforce(w, g.rgrid)
    g.set_fixforce(dr, val, der)
    g.set_fixpot(*fixpot(w, g.rgrid))
    x, y, vx, vy, iz = particles(w, n, seed, g.rgrid)
    g.restrt(x, y, vx, vy, iz)
    g.masset()
    return g


def grid_kwargs3(w: Workload3, n=None):
    """Arguments of P3DGrid / the oracle's P3D for workload w."""
    kw = grid_kwargs(w, n)
    kw["ngz"] = w.ngz
    kw["dzg"] = w.dz_model * w.lscale()
    return kw


def zmax3(w: Workload3):
    """zm = dzg (ngz - 1) / 2 in FP32, as the context computes it."""
    dz = np.float32(w.dz_model * w.lscale())
    return float(np.float32(np.float32(0.5) * dz) * np.float32(w.ngz - 1))


def particles3(w: Workload3, n=None, seed=None, rgrid=None):
    """Internal-unit particles (x, y, z, vx, vy, vz, iz) for workload w."""
    n = w.n if n is None else n
    seed = w.seed if seed is None else seed
    R, phi, vR, vphi = disc.exponential_disc(n, seed, Q=w.Q, md=w.md,
                                             rtrunc=w.rtrunc, rcore=w.rcore)
    z, vz = disc.vertical_sech2(R, seed, z0=w.z0, md=w.md, rtrunc=w.rtrunc)
    rg = w.rgrid_d() if rgrid is None else rgrid
    return disc.to_internal3(R, phi, z, vR, vphi, vz, w.lscale(), w.dt, w.zbr_model,
                             w.nstep, rg, zmax3(w))


def make_grid3(w: Workload3, precision="fp32", n=None, seed=None, green=True):
    """A ready P3DGrid for workload w: Green's tables, the rigid halo as a
    spherical force table, particles loaded and assigned."""
    from .lib15 import P3DGrid
    g = P3DGrid(precision=precision, **grid_kwargs3(w, n))
    if green:
        g.greenm()
    rmax = math.sqrt(g.rgrid ** 2 + g.zm ** 2)
    dr, val, der = fixforce(w, rmax)
    g.set_fixforce(dr, val, der)
    g.restrt(*particles3(w, n, seed, g.rgrid))
    g.masset()
    return g
