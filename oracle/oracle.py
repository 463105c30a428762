"""CPU oracle for GALAXY's 2-D polar-grid particle-mesh step.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` may import this module, and only as the
checker: the product (``galaxy_amd``) never calls it.

The per-particle loops and the field solve live in ``liboracle.so`` (built
from ``orc_p2d.c`` by ``oracle/Makefile``); this module restates the step
schedule around them in numpy, following the reference routines cited on
each method (paths relative to the reference tree).

Precision modes:
  ``"f"``  the reference's own FP32 arithmetic (particles, weights, push and
           zone bookkeeping in float32; field solve in double, rounded);
  ``"d"``  the same formulas promoted to FP64.

Mass sums are kept twice: ``grdmss`` in the reference's sequential order and
arithmetic, and ``mfx`` as order-independent fixed-point int64 (weight x
2^qbits), which is the accumulation the GPU path uses.  ``use_fx`` selects
which one feeds the field solve.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
MAXZ = 10


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


def _load():
    if not os.path.exists(_LIB):
        build()
    return C.CDLL(_LIB)


_lib = _load()
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


class StepCfg(C.Structure):
    _fields_ = [
        ("nr", C.c_int), ("na", C.c_int), ("nsect", C.c_int),
        ("nzones", C.c_int), ("mzone", C.c_int), ("offanal", C.c_int),
        ("fixforce", C.c_int),
        ("nstep", C.c_int * MAXZ),
        ("zbr", C.c_double * MAXZ),
        ("alpha", C.c_double), ("rgrid", C.c_double), ("thmax", C.c_double),
        ("nfr", C.c_int), ("fr_dr", C.c_double),
        ("fr_val", C.c_void_p), ("fr_der", C.c_void_p),
    ]


def _sig(name, res, args):
    fn = getattr(_lib, name)
    fn.restype = res
    fn.argtypes = args
    return fn


_lib.orc_alpha_f.restype = C.c_float
_lib.orc_alpha_f.argtypes = [C.c_int]
_lib.orc_alpha_d.restype = C.c_double
_lib.orc_alpha_d.argtypes = [C.c_int]
_sig("orc_uofgr_f", C.c_float, [C.c_float, C.c_float])
_sig("orc_uofgr_d", C.c_double, [C.c_double, C.c_double])
_sig("orc_grofu_f", C.c_float, [C.c_float, C.c_float])
_sig("orc_grofu_d", C.c_double, [C.c_double, C.c_double])
for _s, _p, _t in (("f", _f32p, C.c_float), ("d", _f64p, C.c_double)):
    _sig("orc_weight_" + _s, None, [C.c_int, _p, _p, C.c_int, C.c_int, _t,
                                    C.c_int, _t, _i32p, _p, _p])
    _sig("orc_massgn_" + _s, None, [C.c_int, _p, _p, _i32p, C.c_int, C.c_int,
                                    _t, C.c_int, C.c_int, _p, _i64p, C.c_int,
                                    _i32p, _i32p])
    _sig("orc_push_" + _s, C.c_int, [C.POINTER(StepCfg), C.c_int, _p, _p, _p,
                                     _p, _i32p, _p, _p, _p, _i64p, C.c_int,
                                     _i32p, _i32p, _u8p])
    _sig("orc_push_par_" + _s, C.c_int, [C.POINTER(StepCfg), C.c_int, _p, _p, _p,
                                         _p, _i32p, _p, _p, _i64p, C.c_int,
                                         _i32p, _i32p, _u8p])
    _sig("orc_tmcent_" + _s, None, [C.POINTER(StepCfg), C.c_int, _p, _p, _p, _p,
                                    _i32p, _p, _p, C.c_int])
_sig("orc_fixforce", C.c_double, [C.c_double, C.c_int, C.c_double, _f64p, _f64p])
for _n in ("orc_rfftf", "orc_rfftb", "orc_cost", "orc_sint"):
    _sig(_n, None, [C.c_int, _f64p, _f64p])
_sig("orc_green", None, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                         C.c_double, C.c_int, _f64p, C.c_void_p])
_sig("orc_fndf", None, [C.c_int, C.c_int, C.c_int, _i32p, _f64p, _f64p,
                        C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _f64p])
_sig("orc_polcat", None, [C.c_int, C.c_int, _f64p])
_sig("orc_polcat_f", None, [C.c_int, C.c_int, _f32p])
_sig("orc_direct_field", None, [C.c_int, C.c_int, C.c_int, _f64p, _f64p, _f64p])


# ---------------------------------------------------------------------------
# thin functional wrappers


def rfftf(x):
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty_like(x)
    _lib.orc_rfftf(len(x), x, y)
    return y


def rfftb(x):
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty_like(x)
    _lib.orc_rfftb(len(x), x, y)
    return y


def cost(x):
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty_like(x)
    _lib.orc_cost(len(x), x, y)
    return y


def sint(x):
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty_like(x)
    _lib.orc_sint(len(x), x, y)
    return y


def alpha(na, prec="d"):
    return _lib.orc_alpha_f(na) if prec == "f" else _lib.orc_alpha_d(na)


def grofu(u, a, prec="d"):
    return _lib.orc_grofu_f(u, a) if prec == "f" else _lib.orc_grofu_d(u, a)


def uofgr(r, a, prec="d"):
    return _lib.orc_uofgr_f(r, a) if prec == "f" else _lib.orc_uofgr_d(r, a)


def weight(x, y, nr, na, prec="d", fold=0, nsect=1):
    """Cell numbers (0-based, = ncl-1), CIC weights (n,4) and radii
    (SRC/lib15/weight.f:61-97)."""
    dt = np.float32 if prec == "f" else np.float64
    x = np.ascontiguousarray(x, dt)
    y = np.ascontiguousarray(y, dt)
    n = len(x)
    cell = np.empty(n, np.int32)
    wt = np.empty((n, 4), dt)
    rr = np.empty(n, dt)
    a = alpha(na, prec)
    thmax = dt(na) / dt(nsect)
    fn = _lib.orc_weight_f if prec == "f" else _lib.orc_weight_d
    fn(n, x, y, nr, na, a, int(fold), thmax, cell, wt, rr)
    return cell, wt, rr


def qbits_for(n: int) -> int:
    """Fixed-point scale exponent: the int64 per-cell sum of n weights in
    [0, 1] stays below 2^62."""
    return 62 - max(1, math.ceil(math.log2(n + 1)))


def green(nr, na, ng, nsect, alpha_d, softl, tsoft, direct=False):
    """Green's tables G[3, ng, nr, nr] (+ direct kernels D[3, nr, nr, mmax])
    (SRC/lib15/p2grnm.f:98-203)."""
    mmax = na // 2 + 1
    G = np.zeros((3, ng, nr, nr), np.float64)
    D = np.zeros((3, nr, nr, mmax), np.float64) if direct else None
    _lib.orc_green(nr, na, ng, nsect, alpha_d, softl, tsoft, G,
                   D.ctypes.data if direct else None)
    return G, D


def fixforce(r, dr, val, der):
    val = np.ascontiguousarray(val, np.float64)
    der = np.ascontiguousarray(der, np.float64)
    return np.array([_lib.orc_fixforce(float(q), len(val), dr, val, der)
                     for q in np.atleast_1d(r)])


def direct_field(nr, na, type_, D, rho):
    out = np.zeros(nr * na, np.float64)
    _lib.orc_direct_field(nr, na, type_, np.ascontiguousarray(D),
                          np.ascontiguousarray(rho, np.float64), out)
    return out


# ---------------------------------------------------------------------------
# full p2d step


class P2D:
    """Oracle state for one 2-D polar-grid model (reference COMMON blocks
    /gridsi/, /zones/, /point/ and arrays grdmss, grdmoz, grdfld)."""

    def __init__(self, nr, na, hash_, softl, tsoft=1, nsect=1, skip=(),
                 nzones=1, nstep=(1,), zbr=(0.0,), pmass=1.0, offanal=True,
                 fixforce=None, prec="d", use_fx=True, qbits=None,
                 green_tables=None):
        self.nr, self.na, self.nsect = nr, na, nsect
        self.mesh = nr * na
        self.mmax = na // 2 + 1
        self.ng = min(hash_ + 1, self.mmax)          # p2dset.f:40, setgrd.f:89
        self.prec = prec
        self.dt = np.float32 if prec == "f" else np.float64
        self.alpha = alpha(na, prec)
        # rgrid = grofu(nr - 1) for jmass = 2 (setgrd.f:127-128)
        self.rgrid = grofu(self.dt(nr - 1), self.alpha, prec)
        self.thmax = self.dt(na) / self.dt(nsect)
        self.softl, self.tsoft = softl, tsoft
        # filter (filset.f:118-141): lg[m] true = harmonic m skipped
        lg = np.ones(max(self.mmax, 1), np.int32)
        lg[0] = 0
        for m in range(1, self.ng):
            lg[m] = 0 if (nsect == 1 or m % nsect == 0) else 1
        for m in skip:
            lg[m] = 1
        self.lg = lg
        self.nzones = nzones
        self.nstep = list(nstep)
        self.zbr = list(zbr)
        self.pmass = pmass
        self.offanal = offanal
        self.fixforce = fixforce   # (dr, val, der) or None
        self.use_fx = use_fx
        self.qbits = qbits
        if green_tables is None:
            G, _ = green(nr, na, self.ng, nsect, alpha(na, "d"), softl, tsoft)
        else:
            G = green_tables
        self.G = np.ascontiguousarray(G)
        self.istep = 0
        self.mzone = nzones
        self.jrad = np.full(nzones, nr, np.int32)
        self.krad = np.zeros(nzones, np.int32)
        self.lstep = np.zeros((2, nzones), np.int64)
        self.mstep = np.zeros(nzones, np.int64)
        self.grdmss = np.zeros((nzones, self.mesh), self.dt)
        self.grdmoz = np.zeros((nzones, self.mesh), self.dt)
        self.mfx = np.zeros((nzones, self.mesh), np.int64)
        self.mozfx = np.zeros((nzones, self.mesh), np.int64)
        self.rho = np.zeros(self.mesh, np.float64)
        self.fld = np.zeros((3, self.mesh), np.float64)
        self.noff = 0

    # -- particles ---------------------------------------------------------
    def load(self, x, y, vx, vy, iz=None):
        """restrt equivalent: particle coordinates in grid units, velocities
        as displacement per step; iz zones (1..nzones, nzones+1 = off)."""
        self.x = np.ascontiguousarray(x, self.dt).copy()
        self.y = np.ascontiguousarray(y, self.dt).copy()
        self.vx = np.ascontiguousarray(vx, self.dt).copy()
        self.vy = np.ascontiguousarray(vy, self.dt).copy()
        n = len(self.x)
        if iz is None:
            iz = np.ones(n, np.int32)
        self.iz = np.ascontiguousarray(iz, np.int32).copy()
        if self.qbits is None:
            self.qbits = qbits_for(n)
        self.noff = int(np.sum(self.iz == self.nzones + 1))

    # -- zone mass bookkeeping ---------------------------------------------
    def _masscl(self, normal):
        """masscl (SRC/lib15/masscl.f:31-146)."""
        for z in range(1, self.mzone + 1):
            ns = self.nstep[z - 1]
            if normal:
                self.mstep[z - 1] = self.istep + ns
            else:
                self.mstep[z - 1] = 0 if self.istep <= 0 else \
                    ((self.istep - 1) // ns + 1) * ns
        for z in range(2, self.mzone + 1):
            self.grdmoz[z - 1] = self.grdmss[z - 1]
            self.mozfx[z - 1] = self.mfx[z - 1]
            self.lstep[0, z - 1] = self.lstep[1, z - 1]
            self.lstep[1, z - 1] = self.mstep[z - 1]
        for z in range(1, self.mzone + 1):
            self.grdmss[z - 1] = 0
            self.mfx[z - 1] = 0
            self.jrad[z - 1] = self.nr
            self.krad[z - 1] = 0

    def _set_mzone(self):
        """mzone = max{i : istep mod nstep(i) == 0} (step.f:135-140)."""
        m = 1
        if self.nzones > 1:
            for i in range(1, self.nzones + 1):
                if self.istep % self.nstep[i - 1] == 0:
                    m = i
        self.mzone = m

    def _zoncmb_scaled(self):
        """zoncmb (SRC/lib15/zoncmb.f:40-125) then scaled (scaled.f:91).
        Reference mode: in REAL arithmetic in the reference's order.
        Fixed-point mode: in double from the integer sums, rounded once."""
        self._set_mzone()
        q = 2.0 ** -self.qbits
        fr = []
        for z in range(2, self.nzones + 1):
            ns = self.nstep[z - 1]
            i = self.lstep[1, z - 1] - self.istep
            if i < 0 or i > ns:
                raise RuntimeError("ZONCMB: new mass array at a useless time")
            fnew = None if i == ns else i
            fold = None
            if i != 0:
                j = self.istep - self.lstep[0, z - 1]
                if self.lstep[0, z - 1] < 0 or j < 0 or j > ns:
                    raise RuntimeError("ZONCMB: old mass array at a useless time")
                fold = j
            fr.append((z, fnew, fold, ns))
        if self.use_fx:
            acc = self.mfx[0].astype(np.float64) * q
            for z, fnew, fold, ns in fr:
                if fnew is not None:
                    acc = acc + (1.0 - fnew / ns) * (self.mfx[z - 1].astype(np.float64) * q)
                if fold is not None:
                    acc = acc + (1.0 - fold / ns) * (self.mozfx[z - 1].astype(np.float64) * q)
            rho = acc * self.pmass
            self.rho = rho.astype(self.dt).astype(np.float64)
        else:
            dt = self.dt
            g1 = self.grdmss[0].copy()
            for z, fnew, fold, ns in fr:
                if fnew is not None:
                    f = dt(1.0) - dt(fnew) / dt(ns)
                    g1 = g1 + f * self.grdmss[z - 1]
                if fold is not None:
                    f = dt(1.0) - dt(fold) / dt(ns)
                    g1 = g1 + f * self.grdmoz[z - 1]
            g1 = g1 * dt(self.pmass)
            self.rho = g1.astype(np.float64)

    def masset(self):
        """masset (SRC/lib15/masset.f): assign every on-grid particle."""
        self.mzone = self.nzones
        self._masscl(False)
        fn = _lib.orc_massgn_f if self.prec == "f" else _lib.orc_massgn_d
        fn(len(self.x), self.x, self.y, self.iz, self.nr, self.na, self.alpha,
           self.nsect, self.nzones, self.grdmss.reshape(-1), self.mfx.reshape(-1),
           self.qbits, self.jrad, self.krad)
        self._zoncmb_scaled()

    # -- field --------------------------------------------------------------
    def source_range(self):
        """Source rings jrs..krs (p2fndf.f:26-36); updates krad like the
        reference."""
        jrs, krs = self.nr, 0
        for ipm in range(1, self.nzones + 1):
            jrs = min(jrs, int(self.jrad[ipm - 1]))
            krs = max(krs, int(self.krad[ipm - 1]))
            if ipm > 1:
                self.krad[ipm - 1] = max(int(self.krad[ipm - 1]), krs)
        if jrs > krs or krs == 0:
            raise RuntimeError("P2FNDF: invalid radial range for source masses")
        return jrs, krs

    def field_range(self, lpot):
        """Field rings jrf..krf (findf.f:42-54, p2fndf.f:42-55)."""
        full = lpot or self.mzone == self.nzones
        jzone, kzone = (0, self.nzones + 1) if full else (1, self.mzone)
        jrf = 1
        if jzone > 0 and self.jrad[jzone - 1] < self.nr:
            jrf = int(self.jrad[jzone - 1])
        krf = self.nr
        if kzone <= self.nzones:
            krf = int(self.krad[kzone - 1])
            while krf < jrf:
                kzone += 1
                if kzone > self.nzones:
                    raise RuntimeError("P2FNDF: kzone > nzones")
                krf = int(self.krad[kzone - 1])
        return jrf, krf

    def findf(self, lpot=False):
        """findf -> p2fndf -> polcat (SRC/lib15/findf.f:27-93)."""
        jrs, krs = self.source_range()
        jrf, krf = self.field_range(lpot)
        ntypes = 3 if lpot else 2
        rho = self.rho.copy()
        _lib.orc_fndf(self.nr, self.na, self.ng, self.lg, self.G.reshape(-1),
                      rho, jrs, krs, jrf, krf, ntypes, self.fld.reshape(-1))
        _lib.orc_polcat(self.nr, self.na, self.fld.reshape(-1))
        self.ranges = (jrs, krs, jrf, krf)
        return self.fld

    # -- step -------------------------------------------------------------
    def stepcfg(self):
        c = StepCfg()
        c.nr, c.na, c.nsect = self.nr, self.na, self.nsect
        c.nzones, c.mzone = self.nzones, self.mzone
        c.offanal = int(bool(self.offanal))
        for i in range(self.nzones):
            c.nstep[i] = self.nstep[i]
            c.zbr[i] = self.zbr[i]
        c.alpha, c.rgrid, c.thmax = float(self.alpha), float(self.rgrid), float(self.thmax)
        if self.fixforce is not None:
            dr, val, der = self.fixforce
            self._fv = np.ascontiguousarray(val, np.float64)
            self._fd = np.ascontiguousarray(der, np.float64)
            c.fixforce, c.nfr, c.fr_dr = 1, len(self._fv), dr
            c.fr_val, c.fr_der = self._fv.ctypes.data, self._fd.ctypes.data
        return c

    def step(self, field=None, parallel=False):
        """step (SRC/lib15/step.f:135-290) for a single grid, no perturber,
        no guard zones.  ``field`` (2, mesh) in the working precision
        overrides the solved Cartesian field (used to inject the GPU's).
        ``parallel``: the particle loop on every host thread (OpenMP); the
        fixed-point sums are the serial loop's exactly, the reference-order
        REAL sums ``grdmss`` are not formed (needs ``use_fx``)."""
        self._set_mzone()
        self._masscl(True)
        if field is None:
            field = self.fld[:2].astype(self.dt)
        fx = np.ascontiguousarray(field[0], self.dt)
        fy = np.ascontiguousarray(field[1], self.dt)
        c = self.stepcfg()
        self.moved = np.zeros(len(self.x), np.uint8)
        if parallel:
            if not self.use_fx:
                raise ValueError("the parallel loop forms the fixed-point sums only")
            fn = _lib.orc_push_par_f if self.prec == "f" else _lib.orc_push_par_d
            dn = fn(C.byref(c), len(self.x), self.x, self.y, self.vx, self.vy,
                    self.iz, fx, fy, self.mfx.reshape(-1), self.qbits, self.jrad,
                    self.krad, self.moved)
        else:
            fn = _lib.orc_push_f if self.prec == "f" else _lib.orc_push_d
            dn = fn(C.byref(c), len(self.x), self.x, self.y, self.vx, self.vy,
                    self.iz, fx, fy, self.grdmss.reshape(-1), self.mfx.reshape(-1),
                    self.qbits, self.jrad, self.krad, self.moved)
        self.noff += dn
        self.istep += 1
        self._zoncmb_scaled()
        return dn

    def integrals(self, fixpot=None, centred=False):
        """icheck's global integrals (SRC/lib15/icheck.f:24-70) in FP64 numpy,
        as galhip_integrals defines them: mass, cm (mass x position),
        momentum, sum m v^2, Lz with velocities per zone-1 step; sum m Phi
        over the grid's particles from the potential of the last
        findf(lpot=True), CIC-interpolated at the particle's own cell and
        weights; sum m Phi_fixed from a (dr, val, der) Hermite table.
        centred: velocities v + a/2 with the current field (anlgrp's
        time-centring before icheck) before the zone step factor."""
        x = self.x.astype(np.float64)
        y = self.y.astype(np.float64)
        iz = self.iz
        zs = np.minimum(iz, self.nzones)
        sf = 1.0 / np.asarray(self.nstep, np.float64)[np.maximum(zs, 1) - 1]
        vx0, vy0 = self.vx, self.vy
        if centred:
            # v + a/2 in the working precision: tmcent's un-centring step
            sv = (self.vx.copy(), self.vy.copy())
            self.tmcent(start=False)
            vx0, vy0 = self.vx, self.vy
            self.vx, self.vy = sv
        vx = vx0.astype(np.float64) * sf
        vy = vy0.astype(np.float64) * sf
        on = iz >= 1
        pm = float(self.pmass)
        out = np.zeros(9)
        out[0] = pm * on.sum()
        out[1], out[2] = pm * x[on].sum(), pm * y[on].sum()
        out[3], out[4] = pm * vx[on].sum(), pm * vy[on].sum()
        out[5] = pm * (vx[on] ** 2 + vy[on] ** 2).sum()
        out[6] = pm * (x[on] * vy[on] - y[on] * vx[on]).sum()
        grid = (iz >= 1) & (iz <= self.nzones)
        cell, w, _ = weight(self.x[grid], self.y[grid], self.nr, self.na, self.prec)
        ring, spoke = cell // self.na, cell % self.na
        s2 = np.where(spoke + 1 == self.na, 0, spoke + 1)
        ic = np.stack([ring * self.na + spoke, ring * self.na + s2,
                       (ring + 1) * self.na + spoke, (ring + 1) * self.na + s2], 1)
        pot = np.append(self.fld[2].astype(np.float64), 0.0)
        ic = np.where(ic < self.mesh, ic, self.mesh)
        out[7] = pm * (w.astype(np.float64) * pot[ic]).sum()
        if fixpot is not None:
            dr, val, der = fixpot
            out[8] = pm * fixforce(np.hypot(x[on], y[on]), dr, val, der).sum()
        return out

    def tmcent(self, start=True, field=None):
        """tmcent (SRC/lib15/tmcent.f, tmcgrp.f:26-36, retarded-velocity
        leap-frog): every particle's velocity minus (start) or plus half its
        getacc acceleration from the current field."""
        if field is None:
            field = self.fld[:2].astype(self.dt)
        fx = np.ascontiguousarray(field[0], self.dt)
        fy = np.ascontiguousarray(field[1], self.dt)
        c = self.stepcfg()
        fn = _lib.orc_tmcent_f if self.prec == "f" else _lib.orc_tmcent_d
        fn(C.byref(c), len(self.x), self.x, self.y, self.vx, self.vy, self.iz,
           fx, fy, int(bool(start)))


# ---------------------------------------------------------------------------
# log-spiral coefficients


def lgspia(x, y, iz, nzones, offanal, nsect, np_, nm, lscale):
    """Log-spiral coefficients of the particles (SRC/lib15/lgspia.f:44-110,
    normalised by the population's weight as phyprp.f:182-197 does): for
    harmonic index im = 1..nm and pitch index j = 1..np_,
        A[j, im] = sum_i exp(i im (nsect theta_i + k_j a_i)) / count,
    a_i = nsect ln(R_i / lscale) / 4, over equal-mass particles on the grid
    (and off it when offanal), in FP64 with the reference's recurrences
    (powers of e^{i nsect theta} and e^{i a} by repeated complex products).
    Pitch index j = jc + k (jc = np_ // 2 + 1, k >= 0) holds pitch k, and
    its mirror np_ + 1 - j pitch -k from the conjugate: k_j = j - jc for odd
    np_; for even np_ index np_ / 2 is never filled and stays zero.  Returns (cos part, sin part),
    each [np_ * nm] in the reference's order (j - 1) nm + (im - 1)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    iz = np.asarray(iz)
    use = (iz >= 1) & (iz <= nzones) if not offanal else (iz >= 1)
    rc = np.sqrt(x * x + y * y)
    use &= rc > 0
    x, y, rc = x[use], y[use], rc[use]
    z1 = (x + 1j * y) / rc
    zs = z1.copy()
    for _ in range(2, nsect + 1):
        zs = zs * z1
    a = 0.25 * nsect * np.log(rc / lscale)
    w1 = np.cos(a) + 1j * np.sin(a)
    A = np.zeros((np_, nm), np.complex128)
    jc = np_ // 2 + 1
    zm = np.ones_like(zs)
    wm = np.ones_like(w1)
    for im in range(1, nm + 1):
        zm = zm * zs
        wm = wm * w1
        A[jc - 1, im - 1] = zm.sum()
        g = np.ones_like(wm)
        for j in range(jc + 1, np_ + 1):
            g = g * wm
            A[j - 1, im - 1] = (g * zm).sum()
            A[np_ - j, im - 1] = (np.conj(g) * zm).sum()
    A /= max(int(use.sum()), 1)
    return A.real.reshape(-1), A.imag.reshape(-1)
