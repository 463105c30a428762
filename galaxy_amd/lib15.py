"""Host-side mirror of GALAXY's lib15 interface for the p2d polar grid.

``P2DGrid`` owns one libgalhip context.  Its methods carry the names of the
reference routines they replace and take the same meaning of arguments
(SRC/lib15/<name>.f):

  ``P2DGrid(...)``  p2dset + setgrd + filset + setspc (grid and zones)
  ``greenm()``      greenm -> p2grnm (Green's tables, built on the GPU in FP64)
  ``restrt()``      restrt, particle part (upload)
  ``unload()``      unload (download in load order)
  ``masset()``      masset
  ``findf(lpot)``   findf -> p2fndf -> polcat
  ``p2fndf(j, k)``  p2fndf with an explicit zone range
  ``step()``        step
  ``run(n)``        galaxy's loop of findf + step
  ``weight()``, ``grdacc()``  group-level routines for parity checks

``P3DGrid`` is the same interface for the 3-D polar grid (p3dset; planes
in z, three coordinates per particle; findf -> p3fndf).

Errors raise ``GalhipError`` where the reference would call crash().
Arrays are numpy; coordinates in grid units, velocities in displacement per
step (SRC/lib15/setrun.f:858-873).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (GALHIP_FP32, GALHIP_FP64, MAXSKIP, MAXZONES, P2DParams,
                   P3DParams, check, load)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class P2DGrid:
    """A 2-D polar grid with its particles on one GPU."""

    def __init__(self, nr, na, hash_, softl, tsoft=1, nsect=1, skip=(),
                 nzones=1, nstep=(1,), zbr=(0.0,), offanal=True, pmass=1.0,
                 precision="fp32"):
        p = self._params(nr, na, hash_, softl, tsoft, nsect, skip, nzones, nstep, zbr,
                         offanal, pmass, precision)
        h = C.c_void_p()
        check(load().galhip_p2dset(C.byref(p), C.byref(h)), "p2dset")
        self._init_handle(h, nr, na, nsect, nzones)

    def _params(self, nr, na, hash_, softl, tsoft, nsect, skip, nzones, nstep, zbr,
                offanal, pmass, precision):
        p = P2DParams()
        p.nr, p.na, p.hash, p.nsect = nr, na, hash_, nsect
        skip = list(skip)
        if len(skip) > MAXSKIP:
            raise ValueError("at most %d skipped harmonics" % MAXSKIP)
        p.nskip = len(skip)
        for i, m in enumerate(skip):
            p.skip[i] = m
        p.softl, p.tsoft = float(softl), int(tsoft)
        if nzones > MAXZONES:
            raise ValueError("at most %d zones" % MAXZONES)
        p.nzones = nzones
        for i in range(nzones):
            p.nstep[i] = int(nstep[i])
            p.zbr[i] = float(zbr[i])
        p.offanal = int(bool(offanal))
        p.pmass = float(pmass)
        if precision not in ("fp32", "fp64"):
            raise ValueError("precision must be 'fp32' or 'fp64'")
        p.precision = GALHIP_FP32 if precision == "fp32" else GALHIP_FP64
        self.dtype = np.float32 if precision == "fp32" else np.float64
        self.precision = precision
        return p

    def _init_handle(self, h, nr, na, nsect, nzones):
        lib = load()
        self._lib, self._h = lib, h
        g = np.zeros(6)
        check(lib.galhip_geometry(h, _ptr(g)), "geometry")
        self.alpha, self.rgrid, self.thmax = g[0], g[1], g[2]
        self.ng, self.mmax, self.mesh = int(g[3]), int(g[4]), int(g[5])
        self.nr, self.na, self.nsect, self.nzones = nr, 2 * (na // 2), nsect, nzones
        self.n = 0

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.galhip_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc, what):
        check(rc, what)

    # -- set-up ---------------------------------------------------------------
    def greenm(self):
        self._c(self._lib.galhip_greenm(self._h), "greenm")

    def get_green(self):
        G = np.zeros((3, self.ng, self.nr, self.nr), np.float64)
        self._c(self._lib.galhip_get_green(self._h, _ptr(G)), "get_green")
        return G

    def set_green(self, G):
        G = np.ascontiguousarray(G, np.float64)
        assert G.shape == (3, self.ng, self.nr, self.nr)
        self._c(self._lib.galhip_set_green(self._h, _ptr(G)), "set_green")

    def set_fixforce(self, dr, val, der):
        val = np.ascontiguousarray(val, np.float64)
        der = np.ascontiguousarray(der, np.float64)
        self._c(self._lib.galhip_set_fixforce(self._h, len(val), float(dr),
                                              _ptr(val), _ptr(der)), "set_fixforce")

    # -- particles ----------------------------------------------------------------
    def restrt(self, x, y, vx, vy, iz=None):
        dt = self.dtype
        arrs = [np.ascontiguousarray(a, dt) for a in (x, y, vx, vy)]
        n = len(arrs[0])
        iz = np.ones(n, np.int32) if iz is None else np.ascontiguousarray(iz, np.int32)
        self._c(self._lib.galhip_restrt(self._h, n, *[_ptr(a) for a in arrs],
                                        _ptr(iz)), "restrt")
        self.n = n

    def unload(self):
        n = self.n
        out = {k: np.empty(n, self.dtype) for k in ("x", "y", "vx", "vy")}
        out["iz"] = np.empty(n, np.int32)
        self._c(self._lib.galhip_unload(self._h, _ptr(out["x"]), _ptr(out["y"]),
                                        _ptr(out["vx"]), _ptr(out["vy"]),
                                        _ptr(out["iz"])), "unload")
        return out

    # -- time bookkeeping ---------------------------------------------------------
    def set_step(self, istep, lstep=None):
        ls = None
        if lstep is not None:
            ls = np.ascontiguousarray(lstep, np.int32).reshape(2, self.nzones)
        self._c(self._lib.galhip_set_step(self._h, int(istep), _ptr(ls)), "set_step")

    def get_step(self):
        i, m = C.c_int32(), C.c_int32()
        ls = np.zeros((2, self.nzones), np.int32)
        self._c(self._lib.galhip_get_step(self._h, C.byref(i), C.byref(m), _ptr(ls)),
                "get_step")
        return i.value, m.value, ls

    # -- the step -----------------------------------------------------------------
    def masset(self):
        self._c(self._lib.galhip_masset(self._h), "masset")

    def findf(self, lpot=False):
        self._c(self._lib.galhip_findf(self._h, int(bool(lpot))), "findf")

    def p2fndf(self, jzone, kzone, lpot=False):
        self._c(self._lib.galhip_p2fndf(self._h, int(jzone), int(kzone),
                                        int(bool(lpot))), "p2fndf")

    def step(self):
        self._c(self._lib.galhip_step(self._h), "step")

    def tmcent(self, start=True):
        """Time-centre (start) or un-centre the velocities (tmcent)."""
        self._c(self._lib.galhip_tmcent(self._h, int(bool(start))), "tmcent")

    def run(self, nsteps):
        self._c(self._lib.galhip_run(self._h, int(nsteps)), "run")

    def sync(self):
        self._c(self._lib.galhip_sync(self._h), "sync")

    def sort(self):
        self._c(self._lib.galhip_sort(self._h), "sort")

    def set_fixpot(self, dr, val, der):
        """Fixed radial potential table (energy diagnostics)."""
        val = np.ascontiguousarray(val, np.float64)
        der = np.ascontiguousarray(der, np.float64)
        self._c(self._lib.galhip_set_fixpot(self._h, len(val), float(dr), _ptr(val),
                                             _ptr(der)), "set_fixpot")

    def integrals(self, centred=True, noalloc=False):
        """icheck's global integrals (see galhip.h): a dict with mass, cm,
        momentum, ke, lz, w_grid (sum m Phi_grid; nan unless the last findf
        formed the potential), w_fixed and the energy ke + w_grid/2 + w_fixed.
        centred: velocities time-centred with the current field, as icheck.
        noalloc: galhip_integrals_noalloc (static device scratch)."""
        out = np.zeros(9, np.float64)
        fn = self._lib.galhip_integrals_noalloc if noalloc else self._lib.galhip_integrals
        self._c(fn(self._h, int(bool(centred)), _ptr(out)), "integrals")
        d = dict(mass=out[0], cm=out[1:3] / out[0] if out[0] else out[1:3],
                 momentum=out[3:5], ke=0.5 * out[5], lz=out[6], w_grid=out[7],
                 w_fixed=out[8])
        d["energy"] = d["ke"] + 0.5 * d["w_grid"] + d["w_fixed"]
        return d

    def lgspia(self, np_, nm, lscale):
        """Log-spiral coefficients of the particles (lgspia, normalised as
        phyprp does): (cos parts, sin parts), each [np_ * nm] in the
        reference's order (j - 1) nm + (im - 1); on an ensemble one pair per
        member."""
        k = int(np_) * int(nm)
        nblk = getattr(self, "nmem", 1)
        out = np.zeros(nblk * 2 * k)
        self._c(self._lib.galhip_lgspia(self._h, int(np_), int(nm), float(lscale), _ptr(out)),
                "lgspia")
        out = out.reshape(nblk, 2, k)
        if nblk == 1:
            return out[0, 0], out[0, 1]
        return [(o[0], o[1]) for o in out]

    def get_order(self):
        """Load-order index of the particle in each storage slot."""
        ids = np.empty(self.n, np.int32)
        self._c(self._lib.galhip_get_order(self._h, _ptr(ids)), "get_order")
        return ids

    def set_pmass(self, pmass):
        """Particle mass in place of the creation parameter (single grids);
        the last field is invalidated."""
        self._c(self._lib.galhip_set_pmass(self._h, float(pmass)), "set_pmass")

    def set_sort_interval(self, k):
        self._c(self._lib.galhip_set_sort_interval(self._h, int(k)), "set_sort_interval")

    def set_graph(self, on=True):
        """Run the steps of run() as replayed HIP graphs."""
        self._c(self._lib.galhip_set_graph(self._h, int(bool(on))), "set_graph")

    def set_contraction(self, mode, variant=None):
        """0 / "exact": FP32 dot products in the synthesis (default);
        1 / "mfma": split-BF16 GEMM on the matrix cores, tile variant tuned
        on the first solve, or fixed by variant (0, 1, 2)."""
        m = 1 if mode in (1, "mfma") else 0
        if m and variant is not None:
            m = 2 + int(variant)
        self._c(self._lib.galhip_set_contraction(self._h, m), "set_contraction")

    def set_compensated(self, on=True):
        """Compensated position updates (FP32): low words of (x, y) carried."""
        self._c(self._lib.galhip_set_compensated(self._h, int(bool(on))), "set_compensated")

    def get_position_lo(self):
        """Low words of (x, y) in load order (zeros when not compensated)."""
        xl = np.zeros(self.n, self.dtype)
        yl = np.zeros(self.n, self.dtype)
        self._c(self._lib.galhip_get_position_lo(self._h, _ptr(xl), _ptr(yl)),
                "get_position_lo")
        return xl, yl

    def get_contraction(self):
        """(mode, tuned variant, per-variant GEMM microseconds)."""
        mode, var = C.c_int32(), C.c_int32()
        us = np.zeros(3)
        self._c(self._lib.galhip_get_contraction(self._h, C.byref(mode), C.byref(var),
                                                 _ptr(us)), "get_contraction")
        return mode.value, var.value, us

    # -- adjoint (one zone, one sector) ------------------------------------------
    def adj_record(self, slot):
        self._c(self._lib.galhip_adj_record(self._h, int(slot)), "adj_record")

    def adj_restore(self, slot):
        self._c(self._lib.galhip_adj_restore(self._h, int(slot)), "adj_restore")

    def adj_set(self, lam):
        """lam: [4][n] FP64 adjoints of x, y, vx, vy in load order."""
        lam = np.ascontiguousarray(lam, np.float64)
        assert lam.shape == (4, self.n)
        self._c(self._lib.galhip_adj_set(self._h, _ptr(lam)), "adj_set")

    def adj_get(self):
        lam = np.zeros((4, self.n), np.float64)
        self._c(self._lib.galhip_adj_get(self._h, _ptr(lam)), "adj_get")
        return lam

    def set_fixforce_dpar(self, dr, val, der):
        val = np.ascontiguousarray(val, np.float64)
        der = np.ascontiguousarray(der, np.float64)
        self._c(self._lib.galhip_set_fixforce_dpar(self._h, len(val), float(dr), _ptr(val),
                                                    _ptr(der)), "set_fixforce_dpar")

    def adj_step(self, gpar=None):
        """One backward step; gpar (2 FP64) accumulates the parameter terms.
        With gpar None the call does not wait for the GPU; the totals stay
        on the device (adj_get_par)."""
        self._c(self._lib.galhip_adj_step(self._h, None if gpar is None else _ptr(gpar)),
                "adj_step")

    def adj_get_par(self):
        """The parameter sums' totals since the adjoints were set (2 FP64)."""
        gpar = np.zeros(2, np.float64)
        self._c(self._lib.galhip_adj_get_par(self._h, _ptr(gpar)), "adj_get_par")
        return gpar

    def adj_seed_a2(self, lscale, r1=1.0, r2=3.0):
        """Adjoints of L = |A_2|^2 over r1 <= r / lscale <= r2, formed on the
        device.  Returns (L, A_2) with A_2 complex."""
        out = np.zeros(4, np.float64)
        self._c(self._lib.galhip_adj_seed_a2(self._h, float(lscale), float(r1), float(r2),
                                             _ptr(out)), "adj_seed_a2")
        return float(out[0]), complex(out[1], out[2])

    def adj_get_profile(self):
        """(scatter ms, gather ms, steps) since the last call (profiling on)."""
        ms = np.zeros(2)
        k = C.c_int64()
        self._c(self._lib.galhip_adj_get_profile(self._h, _ptr(ms), C.byref(k)),
                "adj_get_profile")
        return float(ms[0]), float(ms[1]), k.value

    def adj_field(self, g):
        """K^T: g[2][mesh] -> rbar[mesh] (FP64)."""
        g = np.ascontiguousarray(g, np.float64)
        rb = np.zeros(self.mesh, np.float64)
        self._c(self._lib.galhip_adj_field(self._h, _ptr(g), _ptr(rb)), "adj_field")
        return rb

    # -- tangent-linear (forward mode) ---------------------------------------------
    def tl_set(self, d):
        """Perturbation [4][n] of (x, y, vx, vy), load order, FP64."""
        d = np.ascontiguousarray(d, np.float64)
        assert d.shape == (4, self.n)
        self._c(self._lib.galhip_tl_set(self._h, _ptr(d)), "tl_set")

    def tl_get(self):
        d = np.zeros((4, self.n), np.float64)
        self._c(self._lib.galhip_tl_get(self._h, _ptr(d)), "tl_get")
        return d

    def tl_step(self):
        """Carry the perturbation through the coming step (call after findf,
        before step)."""
        self._c(self._lib.galhip_tl_step(self._h), "tl_step")

    def tl_field(self, rho):
        """K: rho[mesh] -> (F_x, F_y)[2][mesh] (FP64), the last solve's ranges."""
        rho = np.ascontiguousarray(rho, np.float64)
        f = np.zeros((2, self.mesh), np.float64)
        self._c(self._lib.galhip_tl_field(self._h, _ptr(rho), _ptr(f)), "tl_field")
        return f

    # -- group level ---------------------------------------------------------------
    def weight(self, x, y, old=True):
        x = np.ascontiguousarray(x, self.dtype)
        y = np.ascontiguousarray(y, self.dtype)
        n = len(x)
        cell = np.empty(n, np.int32)
        wt = np.empty((n, 4), self.dtype)
        self._c(self._lib.galhip_weight(self._h, n, _ptr(x), _ptr(y), int(bool(old)),
                                        _ptr(cell), _ptr(wt)), "weight")
        return cell, wt

    def grdacc(self, x, y):
        x = np.ascontiguousarray(x, self.dtype)
        y = np.ascontiguousarray(y, self.dtype)
        acc = np.empty((len(x), 2), self.dtype)
        self._c(self._lib.galhip_grdacc(self._h, len(x), _ptr(x), _ptr(y), _ptr(acc)),
                "grdacc")
        return acc

    def getacc(self, x, y, iz):
        """getacc(jst) on a group: acceleration [n][2] at (x, y) for lists iz."""
        x = np.ascontiguousarray(x, self.dtype)
        y = np.ascontiguousarray(y, self.dtype)
        iz = np.ascontiguousarray(iz, np.int32)
        acc = np.empty((len(x), 2), self.dtype)
        self._c(self._lib.galhip_getacc(self._h, len(x), _ptr(x), _ptr(y), _ptr(iz), _ptr(acc)),
                "getacc")
        return acc

    def stpgrp(self, x, y, vx, vy, acc, iz):
        """stpgrp(jst) + offgrd + rezgrp on a group; returns new (x, y, vx, vy, iz)."""
        x, y, vx, vy = (np.array(a, self.dtype) for a in (x, y, vx, vy))
        acc = np.ascontiguousarray(acc, self.dtype)
        iz = np.array(iz, np.int32)
        self._c(self._lib.galhip_stpgrp(self._h, len(x), _ptr(x), _ptr(y), _ptr(vx), _ptr(vy),
                                        _ptr(acc), _ptr(iz)), "stpgrp")
        return x, y, vx, vy, iz

    def massgn(self, x, y, iz):
        """massgn(jst): add a group's mass to the zone sums."""
        x = np.ascontiguousarray(x, self.dtype)
        y = np.ascontiguousarray(y, self.dtype)
        iz = np.ascontiguousarray(iz, np.int32)
        self._c(self._lib.galhip_massgn(self._h, len(x), _ptr(x), _ptr(y), _ptr(iz)), "massgn")

    # -- grids -----------------------------------------------------------------------
    def get_mass(self):
        mfx = np.zeros((self.nzones, self.mesh), np.int64)
        moz = np.zeros((self.nzones, self.mesh), np.int64)
        jr = np.zeros(self.nzones, np.int32)
        kr = np.zeros(self.nzones, np.int32)
        qb = C.c_int32()
        self._c(self._lib.galhip_get_mass(self._h, _ptr(mfx), _ptr(moz), _ptr(jr),
                                          _ptr(kr), C.byref(qb)), "get_mass")
        return mfx, moz, jr, kr, qb.value

    def set_mass(self, mfx=None, moz=None, jrad=None, krad=None):
        a = [None if v is None else np.ascontiguousarray(v, t) for v, t in
             ((mfx, np.int64), (moz, np.int64), (jrad, np.int32), (krad, np.int32))]
        self._c(self._lib.galhip_set_mass(self._h, *[_ptr(v) for v in a]), "set_mass")

    def get_rho(self):
        rho = np.zeros(self.mesh, self.dtype)
        self._c(self._lib.galhip_get_rho(self._h, _ptr(rho)), "get_rho")
        return rho

    def get_field(self):
        f = np.zeros((3, self.mesh), self.dtype)
        self._c(self._lib.galhip_get_field(self._h, _ptr(f)), "get_field")
        return f

    def set_field(self, fld):
        f = np.ascontiguousarray(fld, self.dtype)
        if f.shape[0] == 2:
            f = np.concatenate([f, np.zeros((1, self.mesh), self.dtype)])
        self._c(self._lib.galhip_set_field(self._h, _ptr(f)), "set_field")

    def ranges(self):
        r = np.zeros(4, np.int32)
        self._c(self._lib.galhip_get_ranges(self._h, _ptr(r)), "get_ranges")
        return tuple(int(v) for v in r)

    @property
    def noff(self):
        v = C.c_int64()
        self._c(self._lib.galhip_get_noff(self._h, C.byref(v)), "get_noff")
        return v.value

    # -- profiling ----------------------------------------------------------------
    def set_profile(self, on=True):
        self._c(self._lib.galhip_set_profile(self._h, int(bool(on))), "set_profile")

    def get_profile(self):
        ms, nl, npush = C.c_double(), C.c_int64(), C.c_int64()
        self._c(self._lib.galhip_get_profile(self._h, C.byref(ms), C.byref(nl),
                                             C.byref(npush)), "get_profile")
        return ms.value, nl.value, npush.value


class P3DGrid(P2DGrid):
    """A 3-D polar grid (cylindrical rings x spokes x planes) with its
    particles on one GPU: p3dset + setgrd + filset + setspc.  Zone
    boundaries are spherical radii; the off-grid rule adds |z| > zm."""

    def __init__(self, nr, na, ngz, dzg, hash_, softl, tsoft=1, nsect=1, skip=(),
                 nzones=1, nstep=(1,), zbr=(0.0,), offanal=True, pmass=1.0,
                 precision="fp32", jmass=2):
        lib = load()
        p = P3DParams()
        p.nr, p.na, p.ngz, p.hash, p.nsect = nr, na, ngz, hash_, nsect
        skip = list(skip)
        if len(skip) > MAXSKIP:
            raise ValueError("at most %d skipped harmonics" % MAXSKIP)
        p.nskip = len(skip)
        for i, m in enumerate(skip):
            p.skip[i] = m
        p.dzg, p.softl, p.tsoft = float(dzg), float(softl), int(tsoft)
        if nzones > MAXZONES:
            raise ValueError("at most %d zones" % MAXZONES)
        p.nzones = nzones
        for i in range(nzones):
            p.nstep[i] = int(nstep[i])
            p.zbr[i] = float(zbr[i])
        p.offanal = int(bool(offanal))
        p.pmass = float(pmass)
        p.jmass = int(jmass)
        if precision not in ("fp32", "fp64"):
            raise ValueError("precision must be 'fp32' or 'fp64'")
        p.precision = GALHIP_FP32 if precision == "fp32" else GALHIP_FP64
        self.dtype = np.float32 if precision == "fp32" else np.float64
        self.precision = precision
        self.jmass = int(jmass)
        self.nw = 27 if jmass == 3 else 8
        h = C.c_void_p()
        check(lib.galhip_p3dset(C.byref(p), C.byref(h)), "p3dset")
        self._lib, self._h = lib, h
        g = np.zeros(6)
        check(lib.galhip_geometry(h, _ptr(g)), "geometry")
        self.alpha, self.rgrid, self.thmax = g[0], g[1], g[2]
        self.ng, self.mmax, self.mesh = int(g[3]), int(g[4]), int(g[5])
        self.nr, self.na, self.ngz, self.nsect, self.nzones = nr, 2 * (na // 2), ngz, nsect, nzones
        self.dzg = float(self.dtype(dzg))
        self.zm = float(self.dtype(self.dtype(0.5) * self.dtype(dzg)) *
                        self.dtype(ngz + 1 - self.jmass))
        self.n = 0

    def get_green(self):
        G = np.zeros((4, self.ngz + 1, self.ng, self.nr, self.nr), np.float64)
        self._c(self._lib.galhip_get_green(self._h, _ptr(G)), "get_green")
        return G

    def set_green(self, G):
        G = np.ascontiguousarray(G, np.float64)
        assert G.shape == (4, self.ngz + 1, self.ng, self.nr, self.nr)
        self._c(self._lib.galhip_set_green(self._h, _ptr(G)), "set_green")

    def restrt(self, x, y, z, vx, vy, vz, iz=None):
        dt = self.dtype
        arrs = [np.ascontiguousarray(a, dt) for a in (x, y, z, vx, vy, vz)]
        n = len(arrs[0])
        iz = np.ones(n, np.int32) if iz is None else np.ascontiguousarray(iz, np.int32)
        self._c(self._lib.galhip_restrt3(self._h, n, *[_ptr(a) for a in arrs], _ptr(iz)),
                "restrt3")
        self.n = n

    def unload(self):
        n = self.n
        out = {k: np.empty(n, self.dtype) for k in ("x", "y", "z", "vx", "vy", "vz")}
        out["iz"] = np.empty(n, np.int32)
        self._c(self._lib.galhip_unload3(self._h, *[_ptr(out[k]) for k in
                                                   ("x", "y", "z", "vx", "vy", "vz", "iz")]),
                "unload3")
        return out

    def weight(self, x, y, z, old=True):
        x, y, z = (np.ascontiguousarray(a, self.dtype) for a in (x, y, z))
        n = len(x)
        cell = np.empty(n, np.int32)
        wt = np.empty((n, self.nw), self.dtype)
        self._c(self._lib.galhip_weight3(self._h, n, _ptr(x), _ptr(y), _ptr(z),
                                         int(bool(old)), _ptr(cell), _ptr(wt)), "weight3")
        return cell, wt

    def grdacc(self, x, y, z):
        x, y, z = (np.ascontiguousarray(a, self.dtype) for a in (x, y, z))
        acc = np.empty((len(x), 3), self.dtype)
        self._c(self._lib.galhip_grdacc3(self._h, len(x), _ptr(x), _ptr(y), _ptr(z),
                                         _ptr(acc)), "grdacc3")
        return acc

    def integrals(self, centred=True, noalloc=False):
        """icheck's global integrals on the 3-D grid: mass, cm (3), momentum
        (3), ke, lz, w_grid (nan without a potential), w_fixed, energy.
        noalloc: galhip_integrals_noalloc (static device scratch)."""
        out = np.zeros(11, np.float64)
        fn = self._lib.galhip_integrals_noalloc if noalloc else self._lib.galhip_integrals
        self._c(fn(self._h, int(bool(centred)), _ptr(out)), "integrals")
        d = dict(mass=out[0], cm=out[1:4] / out[0] if out[0] else out[1:4],
                 momentum=out[4:7], ke=0.5 * out[7], lz=out[8], w_grid=out[9],
                 w_fixed=out[10])
        d["energy"] = d["ke"] + 0.5 * d["w_grid"] + d["w_fixed"]
        return d

    def get_field(self):
        f = np.zeros((4, self.mesh), self.dtype)
        self._c(self._lib.galhip_get_field(self._h, _ptr(f)), "get_field")
        return f

    def set_field(self, fld):
        f = np.ascontiguousarray(fld, self.dtype)
        if f.shape[0] == 3:
            f = np.concatenate([f, np.zeros((1, self.mesh), self.dtype)])
        self._c(self._lib.galhip_set_field(self._h, _ptr(f)), "set_field")


class P2DEnsemble(P2DGrid):
    """``nmem`` independent p2d discs on one grid geometry (galhip_ensset).

    Members share the grid, filter, zones, softening, Green's table, fixed
    force and potential tables and the step counter; each has its own
    particles and particle mass (``pmass``, one per member).  greenm, masset,
    findf, step, tmcent, run, sort and the profiling calls act on every
    member; particles, grids and integrals are per member.  No counterpart in
    the reference, which runs one model per process.
    """

    def __init__(self, nmem, nr, na, hash_, softl, pmass, tsoft=1, nsect=1, skip=(),
                 nzones=1, nstep=(1,), zbr=(0.0,), offanal=True, precision="fp32"):
        pm = np.ascontiguousarray(pmass, np.float64)
        if pm.shape != (nmem,):
            raise ValueError("one particle mass per member")
        p = self._params(nr, na, hash_, softl, tsoft, nsect, skip, nzones, nstep, zbr,
                         offanal, 1.0, precision)
        h = C.c_void_p()
        check(load().galhip_ensset(C.byref(p), int(nmem), _ptr(pm), C.byref(h)), "ensset")
        self._init_handle(h, nr, na, nsect, nzones)
        self.nmem = int(nmem)
        self.counts = np.zeros(self.nmem, np.int64)

    def restrt(self, members):
        """members: one (x, y, vx, vy[, iz]) tuple per member."""
        if len(members) != self.nmem:
            raise ValueError("one particle set per member")
        dt = self.dtype
        cols = [[], [], [], [], []]
        counts = []
        for mem in members:
            x = np.ascontiguousarray(mem[0], dt)
            n = len(x)
            counts.append(n)
            cols[0].append(x)
            for k in (1, 2, 3):
                cols[k].append(np.ascontiguousarray(mem[k], dt))
            iz = mem[4] if len(mem) > 4 and mem[4] is not None else np.ones(n, np.int32)
            cols[4].append(np.ascontiguousarray(iz, np.int32))
        arrs = [np.concatenate(c) if c else np.zeros(0) for c in cols]
        arrs = [a.astype(dt) for a in arrs[:4]] + [arrs[4].astype(np.int32)]
        self.counts = np.asarray(counts, np.int64)
        self._c(self._lib.galhip_ens_restrt(self._h, _ptr(self.counts),
                                            *[_ptr(a) for a in arrs]), "ens_restrt")
        self.n = int(self.counts.sum())

    def unload(self):
        """A list of per-member dicts x, y, vx, vy, iz."""
        n = self.n
        out = {k: np.empty(n, self.dtype) for k in ("x", "y", "vx", "vy")}
        out["iz"] = np.empty(n, np.int32)
        self._c(self._lib.galhip_ens_unload(self._h, _ptr(out["x"]), _ptr(out["y"]),
                                            _ptr(out["vx"]), _ptr(out["vy"]),
                                            _ptr(out["iz"])), "ens_unload")
        off = np.concatenate([[0], np.cumsum(self.counts)])
        return [{k: v[off[m]:off[m + 1]] for k, v in out.items()} for m in range(self.nmem)]

    def get_mass(self, member):
        mfx = np.zeros((self.nzones, self.mesh), np.int64)
        moz = np.zeros((self.nzones, self.mesh), np.int64)
        jr = np.zeros(self.nzones, np.int32)
        kr = np.zeros(self.nzones, np.int32)
        qb = C.c_int32()
        self._c(self._lib.galhip_ens_get_mass(self._h, int(member), _ptr(mfx), _ptr(moz),
                                              _ptr(jr), _ptr(kr), C.byref(qb)), "ens_get_mass")
        return mfx, moz, jr, kr, qb.value

    def get_field(self, member):
        f = np.zeros((3, self.mesh), self.dtype)
        self._c(self._lib.galhip_ens_get_field(self._h, int(member), _ptr(f)), "ens_get_field")
        return f

    def set_field(self, member, fld):
        f = np.ascontiguousarray(fld, self.dtype)
        if f.shape[0] == 2:
            f = np.concatenate([f, np.zeros((1, self.mesh), self.dtype)])
        self._c(self._lib.galhip_ens_set_field(self._h, int(member), _ptr(f)), "ens_set_field")

    def ranges(self, member):
        r = np.zeros(4, np.int32)
        self._c(self._lib.galhip_ens_get_ranges(self._h, int(member), _ptr(r)), "ens_get_ranges")
        return tuple(int(v) for v in r)

    @property
    def noff(self):
        v = np.zeros(self.nmem, np.int64)
        self._c(self._lib.galhip_ens_get_noff(self._h, _ptr(v)), "ens_get_noff")
        return v

    def integrals(self, centred=True):
        """P2DGrid.integrals for every member: a list of dicts."""
        out = np.zeros((self.nmem, 9), np.float64)
        self._c(self._lib.galhip_ens_integrals(self._h, int(bool(centred)), _ptr(out)),
                "ens_integrals")
        res = []
        for o in out:
            d = dict(mass=o[0], cm=o[1:3] / o[0] if o[0] else o[1:3], momentum=o[3:5],
                     ke=0.5 * o[5], lz=o[6], w_grid=o[7], w_fixed=o[8])
            d["energy"] = d["ke"] + 0.5 * d["w_grid"] + d["w_fixed"]
            res.append(d)
        return res

    def get_rho(self):
        raise NotImplementedError("per-member rho is not exported; use get_field")

    def set_mass(self, *a, **k):
        raise NotImplementedError("set_mass is per grid; not available on an ensemble")

    def grdacc(self, x, y):
        raise NotImplementedError("an ensemble has one field per member")
