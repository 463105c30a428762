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
    (assumed; the workload names neither);
  - ensemble: member counts and Q drawn with numpy's default_rng(4), member
    i seeded i, m <= 8, one zone, dt = 0.05 (assumed).
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



@dataclass
class EnsembleWorkload(Workload):
    """nmem independent discs on one grid: member i has count_i particles
    with log10(count_i) uniform in [log10 nmin, log10 nmax], Toomre Q_i
    uniform in [qmin, qmax] (both drawn with draw_seed) and its own seed
    seed + i.  n is the total particle count."""
    nmem: int = 64
    nmin: float = 1e5
    nmax: float = 4e6
    qmin: float = 1.0
    qmax: float = 2.0
    draw_seed: int = 4

    def members(self, scale=1.0):
        """(counts, Q) of the members; scale multiplies the counts (smaller
        ensembles of the same shape for tests)."""
        rng = np.random.default_rng(self.draw_seed)
        logn = rng.uniform(math.log10(self.nmin), math.log10(self.nmax), self.nmem)
        Q = rng.uniform(self.qmin, self.qmax, self.nmem)
        counts = np.maximum(np.rint(scale * 10.0 ** logn), 0).astype(np.int64)
        return counts, Q

    def member(self, i, n):
        """Workload of member i as a lone disc of n particles."""
        counts, Q = self.members()
        return Workload(self.name + "-m%d" % i, nr=self.nr, na=self.na, hash=self.hash,
                        n=int(n), steps=self.steps, dt=self.dt, nstep=self.nstep,
                        zbr_model=self.zbr_model, eps_model=self.eps_model, Q=float(Q[i]),
                        seed=self.seed + i, md=self.md, rtrunc=self.rtrunc, rcore=self.rcore,
                        redge=self.redge)


# 64 discs on 128 x 256 grids, N_i log-uniform in [1e5, 4e6], Q_i uniform in
# [1, 2], m <= 8 (assumed), one zone and dt = 0.05 (assumed), 300 steps
C4 = EnsembleWorkload("ens-64x(128x256)", nr=128, na=256, hash=8, n=0, steps=300,
                      dt=0.05, seed=0)
C4.n = int(C4.members()[0].sum())

# differentiated run: 128 x 256 grid, N = 1e6, Q = 1.2, 200 steps of dt =
# 0.05, m <= 8 (assumed), one zone; loss |A_2|^2 over R = 1..3 at the end
C5 = Workload("p2d-128x256-n1e6-adjoint", nr=128, na=256, hash=8, n=1_000_000,
              steps=200, dt=0.05, Q=1.2, seed=5)

WORKLOADS = {"C1": C1, "C2": C2, "C3": C3, "C4": C4, "C5": C5}


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


def make_grid(w: Workload, precision="fp32", n=None, seed=None, green=True, compensated=False):
    """A ready P2DGrid for workload w: Green's tables, halo table, particles
    loaded and assigned (masset); compensated: compensated position updates."""
    from .lib15 import P2DGrid
    g = P2DGrid(precision=precision, **grid_kwargs(w, n))
    if compensated:
        g.set_compensated(True)
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


def _member_particles(args):
    w, i, n, rgrid = args
    if n <= 0:
        e = np.zeros(0)
        return (e, e, e, e, np.zeros(0, np.int32))
    return particles(w.member(i, n), n, None, rgrid)


def ensemble_particles(w: EnsembleWorkload, counts, rgrid, workers=None,
                       min_parallel=2_000_000):
    """The members' particles (x, y, vx, vy, iz per member), drawn in
    parallel worker processes (each member from its own seed, so the result
    does not depend on the worker count)."""
    import concurrent.futures as cf
    import os
    jobs = [(w, i, int(n), rgrid) for i, n in enumerate(counts)]
    workers = workers or min(8, os.cpu_count() or 1)
    if workers <= 1 or sum(counts) < min_parallel:
        return [_member_particles(j) for j in jobs]
    with cf.ProcessPoolExecutor(workers) as ex:
        return list(ex.map(_member_particles, jobs))


def make_ensemble(w: EnsembleWorkload, precision="fp32", scale=1.0, green=True, parts=None):
    """A ready P2DEnsemble for workload w (member counts times scale): one
    Green's table, the halo tables, every member's particles (parts, or
    drawn here) loaded and assigned."""
    from .lib15 import P2DEnsemble
    counts, _ = w.members(scale)
    kw = grid_kwargs(w, 1)
    kw.pop("pmass")
    ls = w.lscale()
    pmass = [w.md * ls ** 3 * w.dt * w.dt / max(int(n), 1) for n in counts]
    g = P2DEnsemble(w.nmem, pmass=pmass, precision=precision, **kw)
    if green:
        g.greenm()
    g.set_fixforce(*fixforce(w, g.rgrid))
    g.set_fixpot(*fixpot(w, g.rgrid))
    g.restrt(parts if parts is not None else ensemble_particles(w, counts, g.rgrid))
    g.masset()
    return g
