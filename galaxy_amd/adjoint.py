"""Gradients of a p2d run by the adjoint kernels (galhip_adj_*), with
checkpoint/recompute, and a torch.autograd.Function around them.

No counterpart in the reference.  The differentiated run is galaxy's loop
(findf, step) for one zone and one sector: from initial internal-unit
positions and leap-frog velocities (x0, v0) to the final positions.  The
loss used here is |A_2|^2, the squared m = 2 Fourier coefficient of the
particles between R = 1 and 3 disc scale lengths at the final step
(A_2 = sum e^{-2 i phi} / count over that annulus, equal masses).

Backward pass, for checkpoints every `every` steps: each segment is
recomputed from its checkpoint, recording the state before each step in a
device slot (particles, zone mass sums, ring ranges); then, from the
segment's last step down, the state is restored, findf rebuilds that
step's field bit for bit, and galhip_adj_step takes the adjoints one step
back.  Parameter terms come out of the same pass: gpar[0] (the
derivative for a factor on the particle mass, hence the grid force) and
gpar[1] (the fixed force's parameter derivative table).
"""
from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import configs, disc


def a2_loss(x, y, lscale, r1=1.0, r2=3.0, sel=None):
    """L = |A_2|^2 over r1 <= R <= r2 (R = r / lscale) and dL/dx, dL/dy
    (membership of the annulus held fixed), FP64.  sel: a given membership
    mask instead (a finite-difference check of that derivative at large N,
    where single particles crossing the annulus edges make L jump).

    With c = cos 2 phi = (x^2 - y^2)/r^2, s = sin 2 phi = 2xy/r^2 and
    A = (sum c - i sum s)/count = Ar + i Ai: L = Ar^2 + Ai^2 and
    dL/dphi_j = -4 (Ar s_j + Ai c_j)/count; dphi/dx = -y/r^2, dphi/dy = x/r^2."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    r2_ = x * x + y * y
    if sel is None:
        R = np.sqrt(r2_) / lscale
        sel = (R >= r1) & (R <= r2)
    cnt = max(int(sel.sum()), 1)
    with np.errstate(invalid="ignore", divide="ignore"):
        inv = np.where(r2_ > 0, 1.0 / r2_, 0.0)
    c2 = (x * x - y * y) * inv
    s2 = 2.0 * x * y * inv
    ar = c2[sel].sum() / cnt
    ai = -s2[sel].sum() / cnt
    L = ar * ar + ai * ai
    dphi = (-4.0 / cnt) * (ar * s2 + ai * c2)
    dphi[~sel] = 0.0
    dphi *= inv
    return L, -dphi * y, dphi * x


class Trajectory:
    """A p2d grid run for `nsteps` steps with checkpoints every `every`
    steps, and its adjoint.  grid: a P2DGrid with one zone and one sector,
    Green's tables and fixed force set."""

    def __init__(self, grid, nsteps, every=20):
        self.g, self.nsteps, self.every = grid, int(nsteps), int(every)
        self.nck = (self.nsteps + self.every - 1) // self.every
        self.seconds = dict(forward=0.0, recompute=0.0, backward=0.0)

    def forward(self, x, y, vx, vy, iz, unload=True):
        """Load, run, return the final state (dict in load order; None with
        unload False: the state stays on the device)."""
        g = self.g
        g.set_sort_interval(10)
        g.set_step(0)
        g.restrt(x, y, vx, vy, iz)
        g.masset()
        g.sync()
        t0 = time.perf_counter()
        for s in range(self.nsteps):
            if s % self.every == 0:
                g.adj_record(s // self.every)
            g.run(1)
        g.sync()
        self.seconds["forward"] += time.perf_counter() - t0
        return g.unload() if unload else None

    def backward(self, lam=None):
        """lam [4][n]: adjoints of the final (x, y, vx, vy); None when they
        were set on the device after forward (P2DGrid.adj_seed_a2).  Returns
        the adjoints of the initial state and gpar [2]."""
        g, E = self.g, self.every
        if lam is not None:
            g.adj_set(lam)
        base = self.nck
        g.sync()
        for k in range(self.nck - 1, -1, -1):
            m = min(E, self.nsteps - k * E)
            t0 = time.perf_counter()
            g.adj_restore(k)
            for j in range(m):
                g.adj_record(base + j)
                g.run(1)
            g.sync()
            t1 = time.perf_counter()
            for j in range(m - 1, -1, -1):
                g.adj_restore(base + j)
                g.findf(False)
                g.adj_step()
            g.sync()
            t2 = time.perf_counter()
            self.seconds["recompute"] += t1 - t0
            self.seconds["backward"] += t2 - t1
        return g.adj_get(), g.adj_get_par()


    def tangent(self, x, y, vx, vy, iz, d0):
        """FP64 tangent-linear run (galhip_tl_*): the forward run from the
        same state with the perturbation d0 [4][n] (load order) carried
        along step by step; re-sorted every 10 steps as the forward run is.
        Returns the final perturbation [4][n]; the final state stays on the
        device."""
        g = self.g
        g.set_sort_interval(0)
        g.set_step(0)
        g.restrt(x, y, vx, vy, iz)
        g.masset()
        g.tl_set(d0)
        t0 = time.perf_counter()
        for s in range(self.nsteps):
            g.findf(False)
            g.tl_step()
            g.step()
            if (s + 1) % 10 == 0:
                g.sort()
        g.sync()
        self.seconds["tangent"] = self.seconds.get("tangent", 0.0) + time.perf_counter() - t0
        return g.tl_get()


class DiscModel:
    """The differentiated workload: a seeded exponential disc whose initial
    velocities depend on Toomre Q and the disc mass md (sigma_R = Q md
    k_sig(R)), in a rigid halo that keeps the rotation curve fixed, so the
    halo force depends on md too (F_halo = -v_c^2/R - md F_disc1(R))."""

    def __init__(self, w: configs.Workload, precision="fp32", n=None):
        from .lib15 import P2DGrid
        self.w = w
        self.n = w.n if n is None else int(n)
        self.precision = precision
        self.ls = w.lscale()
        kw = configs.grid_kwargs(w, self.n)
        self._kw = kw
        g = P2DGrid(precision=precision, **kw)
        g.greenm()
        self.G = g.get_green()
        self.rgrid = float(g.rgrid)
        g.close()
        R, phi, g1, g2 = disc.disc_draws(self.n, w.seed, w.rtrunc)
        self.R, self.phi, self.g1, self.g2 = R, phi, g1, g2
        self.k_sig, self.k_phi, self.vc2, self.c = disc.velocity_terms(R, w.rtrunc, w.rcore)
        self.cp, self.sp = np.cos(phi), np.sin(phi)
        self.x0 = R * self.cp * self.ls
        self.y0 = R * self.sp * self.ls
        self.iz0 = np.where(R * self.ls > self.rgrid, 2, 1).astype(np.int32)

    def velocities(self, Q, md):
        """Internal-unit initial velocities (numpy FP64) for Q, md."""
        return ic_velocities(self.k_sig, self.k_phi, self.vc2, self.c, self.g1, self.g2,
                             self.cp, self.sp, Q, md, self.ls * self.w.dt)

    def grid(self, md):
        """A fresh grid for disc mass md: particle mass and halo force scale
        with it; the parameter derivative table is dF_halo/dmd."""
        from .lib15 import P2DGrid
        w, ls, ts = self.w, self.ls, self.w.dt
        kw = dict(self._kw)
        kw["pmass"] = md * ls ** 3 * ts * ts / self.n
        g = P2DGrid(precision=self.precision, **kw)
        g.set_green(self.G)
        g.set_fixforce(*disc.halo_table(self.rgrid, ls, ts, md=md, rcore=w.rcore))
        dr, v1, d1 = disc.halo_table(self.rgrid, ls, ts, md=1.0, rcore=w.rcore)
        _, v0, d0 = disc.halo_table(self.rgrid, ls, ts, md=0.0, rcore=w.rcore)
        g.set_fixforce_dpar(dr, v1 - v0, d1 - d0)
        return g

    def _grid_for(self, md):
        """The model's own grid, kept between gradient calls (its device
        buffers and adjoint slots are allocated once), set to disc mass md."""
        g = getattr(self, "_g", None)
        if g is None:
            g = self._g = self.grid(md)
        elif md != self._g_md:
            w, ls, ts = self.w, self.ls, self.w.dt
            g.set_pmass(md * ls ** 3 * ts * ts / self.n)
            g.set_fixforce(*disc.halo_table(self.rgrid, ls, ts, md=md, rcore=w.rcore))
        self._g_md = md
        return g

    def close(self):
        """Frees the grid the gradient calls keep."""
        if getattr(self, "_g", None) is not None:
            self._g.close()
            self._g = None

    def loss(self, Q, md, nsteps):
        """Forward only: L for (Q, md) (no checkpoints)."""
        vx, vy = self.velocities(Q, md)
        return self.loss_v(vx, vy, md, nsteps)

    def loss_v(self, vx, vy, md, nsteps, sel=None):
        g = self.grid(md)
        g.restrt(self.x0, self.y0, vx, vy, self.iz0)
        g.set_sort_interval(10)
        g.masset()
        g.run(int(nsteps))
        out = g.unload()
        g.close()
        return a2_loss(out["x"], out["y"], self.ls, sel=sel)[0]

    def gradient(self, Q, md, nsteps, every=20, profile=False, with_sel=False):
        """L and its gradient: dL/dvx0, dL/dvy0 (internal units), dL/dQ,
        dL/dmd, through the adjoint kernels.  The loss and its adjoint seed
        are formed on the device (galhip_adj_seed_a2); d(v0)/dQ is computed
        on the host in a worker thread while the GPU runs.  profile: HIP
        events around the adjoint particle kernels (result 'kernel_ms':
        scatter, gather).  with_sel: also return the final state's annulus
        membership (load order; 'sel')."""
        t0 = time.perf_counter()
        vx, vy = self.velocities(Q, md)
        with ThreadPoolExecutor(max_workers=1) as pool:
            dv = pool.submit(ic_velocity_dq, self.k_sig, self.k_phi, self.vc2, self.c,
                             self.g1, self.g2, self.cp, self.sp, Q, md, self.ls * self.w.dt)
            g = self._grid_for(md)
            if profile:
                g.set_profile(True)
            tr = Trajectory(g, nsteps, every)
            tr.seconds["setup"] = time.perf_counter() - t0
            tr.forward(self.x0, self.y0, vx, vy, self.iz0, unload=False)
            sel = None
            if with_sel:
                out = g.unload()
                R = np.hypot(out["x"].astype(np.float64), out["y"].astype(np.float64)) / self.ls
                sel = (R >= 1.0) & (R <= 3.0)
            L, _ = g.adj_seed_a2(self.ls, 1.0, 3.0)
            lam0, gpar = tr.backward()
            kms = None
            if profile:
                kms = g.adj_get_profile()
                g.set_profile(False)
            t1 = time.perf_counter()
            dvx, dvy = dv.result()
        lvx, lvy = lam0[2], lam0[3]
        dQ = float(np.dot(lvx, dvx) + np.dot(lvy, dvy))
        dmd = dQ * Q / md + gpar[0] / md + gpar[1]
        tr.seconds["finish"] = time.perf_counter() - t1
        return dict(loss=L, dvx=lvx, dvy=lvy, dQ=dQ, dmd=dmd, gpar=gpar,
                    seconds=dict(tr.seconds), sel=sel, kernel_ms=kms)


    def tangent(self, Q, md, nsteps, d0):
        """Directional derivative of L along the initial-state perturbation
        d0 [4][n] ((x, y, vx, vy), load order, internal units) by the FP64
        tangent-linear run: dL = <dL/dx_N, dx_N>, the seed dL/dx_N formed
        on the device as the adjoint's (galhip_adj_seed_a2).  Returns
        (dL, L, the final perturbation)."""
        vx, vy = self.velocities(Q, md)
        g = self._grid_for(md)
        tr = Trajectory(g, nsteps)
        d = tr.tangent(self.x0, self.y0, vx, vy, self.iz0, d0)
        L, _ = g.adj_seed_a2(self.ls, 1.0, 3.0)
        lam = g.adj_get()
        dL = float(np.dot(lam[0], d[0]) + np.dot(lam[1], d[1]))
        return dL, L, d


def ic_velocities(k_sig, k_phi, vc2, c, g1, g2, cp, sp, Q, md, f):
    """Internal-unit initial velocities of the exponential disc (the
    formulas of disc.exponential_disc on its draws and terms): sigma_R =
    Q md k_sig, vR = sigma_R g1, vphi = sqrt(max(vc2 + sigma_R^2 c, 0)) +
    sigma_R k_phi g2, rotated by (cp, sp) and scaled by f = lscale ts."""
    sR = Q * md * k_sig
    vm = np.sqrt(np.maximum(vc2 + sR * sR * c, 0.0))
    vR = sR * g1
    vphi = vm + sR * k_phi * g2
    return (vR * cp - vphi * sp) * f, (vR * sp + vphi * cp) * f


def ic_velocity_dq(k_sig, k_phi, vc2, c, g1, g2, cp, sp, Q, md, f):
    """d/dQ of ic_velocities (d/dmd is Q/md times it)."""
    sR = Q * md * k_sig
    vphi2 = vc2 + sR * sR * c
    vm = np.sqrt(np.maximum(vphi2, 0.0))
    with np.errstate(divide="ignore", invalid="ignore"):
        dvm = np.where(vphi2 > 0, sR * c / np.where(vm > 0, vm, 1.0), 0.0)
    dvphi = dvm + k_phi * g2           # d/d sigma_R
    dvx = (g1 * cp - dvphi * sp) * f
    dvy = (g1 * sp + dvphi * cp) * f
    ds = sR / Q                        # d sigma_R / dQ
    return dvx * ds, dvy * ds


def torch_function():
    """A torch.autograd.Function mapping (vx0, vy0, md) to the final
    (x, y) of a DiscModel run, its backward the adjoint pass."""
    import torch

    class PMRun(torch.autograd.Function):
        @staticmethod
        def forward(ctx, vx0, vy0, md, model, nsteps, every):
            g = model.grid(float(md))
            tr = Trajectory(g, nsteps, every)
            out = tr.forward(model.x0, model.y0, vx0.detach().cpu().numpy(),
                             vy0.detach().cpu().numpy(), model.iz0)
            ctx.tr, ctx.g, ctx.md = tr, g, float(md)
            return (torch.from_numpy(out["x"].astype(np.float64)),
                    torch.from_numpy(out["y"].astype(np.float64)))

        @staticmethod
        def backward(ctx, gx, gy):
            n = gx.shape[0]
            lam = np.zeros((4, n))
            lam[0] = gx.detach().cpu().numpy()
            lam[1] = gy.detach().cpu().numpy()
            lam0, gpar = ctx.tr.backward(lam)
            ctx.g.close()
            dmd = gpar[0] / ctx.md + gpar[1]
            return (torch.from_numpy(lam0[2]), torch.from_numpy(lam0[3]),
                    torch.tensor(dmd, dtype=torch.float64), None, None, None)

    return PMRun


def torch_loss(model, Q, md, nsteps, every=20):
    """|A_2|^2 as a torch graph: initial velocities from (Q, md) in torch,
    the run through the adjoint-backed Function, A_2 from the final
    positions.  Q and md are 0-d FP64 tensors (requires_grad for their
    gradients)."""
    import torch
    t = lambda a: torch.from_numpy(np.asarray(a, np.float64))
    k_sig, k_phi, vc2, c = t(model.k_sig), t(model.k_phi), t(model.vc2), t(model.c)
    g1, g2, phi = t(model.g1), t(model.g2), t(model.phi)
    sR = Q * md * k_sig
    vm = torch.sqrt(torch.clamp(vc2 + sR * sR * c, min=1e-300))
    vR = sR * g1
    vphi = vm + sR * k_phi * g2
    f = model.ls * model.w.dt
    vx = (vR * torch.cos(phi) - vphi * torch.sin(phi)) * f
    vy = (vR * torch.sin(phi) + vphi * torch.cos(phi)) * f
    x, y = torch_function().apply(vx, vy, md, model, nsteps, every)
    R = torch.sqrt(x * x + y * y).detach() / model.ls
    sel = (R >= 1.0) & (R <= 3.0)
    ph = torch.atan2(y[sel], x[sel])
    cnt = max(int(sel.sum()), 1)
    ar = torch.cos(2 * ph).sum() / cnt
    ai = torch.sin(2 * ph).sum() / cnt
    return ar * ar + ai * ai
