"""CPU oracle for GALAXY's 3-D polar-grid (p3d) particle-mesh step.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` may import this module, and only as the
checker: the product (``galaxy_amd``) never calls it.

The particle loops, Green's tables and field solve are the C restatement in
``orc_p3d.c`` (built into ``liboracle.so`` with the p2d oracle); the zone
and time bookkeeping (masscl, zoncmb, scaled, the radial ranges of
p3fndf.f:44-75) is the p2d oracle's, which the reference shares between the
two grids.  Precision modes as for the p2d oracle: ``"f"`` the reference's
FP32 particle arithmetic, ``"d"`` the same promoted to FP64; field solves in
double.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from oracle import oracle as _o

_lib = _o._lib
_f32p, _f64p, _i32p, _i64p, _u8p = _o._f32p, _o._f64p, _o._i32p, _o._i64p, _o._u8p
MAXZ = 10


class Cfg3(C.Structure):
    _fields_ = [
        ("nr", C.c_int), ("na", C.c_int), ("ngz", C.c_int), ("nsect", C.c_int),
        ("nzones", C.c_int), ("mzone", C.c_int), ("offanal", C.c_int),
        ("fixforce", C.c_int), ("jmass", C.c_int),
        ("nstep", C.c_int * MAXZ),
        ("zbr", C.c_double * MAXZ),
        ("alpha", C.c_double), ("rgrid", C.c_double), ("thmax", C.c_double),
        ("zm", C.c_double), ("dzg", C.c_double),
        ("nfr", C.c_int), ("fr_dr", C.c_double),
        ("fr_val", C.c_void_p), ("fr_der", C.c_void_p),
    ]


_CP = C.POINTER(Cfg3)
for _s, _p in (("f", _f32p), ("d", _f64p)):
    _o._sig("orc3_weight_" + _s, None, [_CP, C.c_int, _p, _p, _p, C.c_int, _i32p, _p, _p])
    _o._sig("orc3_massgn_" + _s, None, [_CP, C.c_int, _p, _p, _p, _i32p, _p, _i64p,
                                        C.c_int, _i32p, _i32p])
    _o._sig("orc3_getacc_" + _s, None, [_CP, C.c_int, _p, _p, _p, _i32p, _p, _p])
    _o._sig("orc3_push_" + _s, C.c_int, [_CP, C.c_int, _p, _p, _p, _p, _p, _p, _i32p,
                                         _p, _p, _i64p, C.c_int, _i32p, _i32p, _u8p])
    _o._sig("orc3_tmcent_" + _s, None, [_CP, C.c_int, _p, _p, _p, _p, _p, _p, _i32p,
                                        _p, C.c_int])
_o._sig("orc3_green", None, [C.c_int, C.c_int, C.c_int, C.c_int, _i32p, C.c_double,
                             C.c_double, C.c_double, C.c_int, _f64p, C.c_void_p])
_o._sig("orc3_fndf", None, [C.c_int, C.c_int, C.c_int, C.c_int, _i32p, _f64p, _f64p,
                            C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _f64p])
_o._sig("orc3_direct_field", None, [C.c_int, C.c_int, C.c_int, C.c_int, _f64p, _f64p,
                                    _f64p])


def filter_flags(na, ng, nsect=1, skip=()):
    """lg[m] true = harmonic m skipped (filset.f:118-141)."""
    mmax = na // 2 + 1
    lg = np.ones(mmax, np.int32)
    lg[0] = 0
    for m in range(1, ng):
        lg[m] = 0 if (nsect == 1 or m % nsect == 0) else 1
    for m in skip:
        lg[m] = 1
    return lg


def green3(nr, na, ngz, ng, lg, dzg, softl, tsoft, direct=False):
    """Tables G[4, ngz+1, ng, nr(source), nr(field)] and, if direct, the
    direct kernels D[4, nr, nr, ngz+1, mmax] (SRC/lib15/p3grnm.f:93-263)."""
    mmax = na // 2 + 1
    G = np.zeros((4, ngz + 1, ng, nr, nr), np.float64)
    D = np.zeros((4, nr, nr, ngz + 1, mmax), np.float64) if direct else None
    _lib.orc3_green(nr, na, ngz, ng, np.ascontiguousarray(lg, np.int32),
                    _o.alpha(na, "d"), dzg, softl, tsoft, G,
                    D.ctypes.data if direct else None)
    return G, D


def fndf3(nr, na, ngz, ng, lg, G, rho, jrs, krs, jrf, krf, ntypes, fld=None):
    """p3manl + p3conv + p3fsyn on the combined mass rho; returns fld[4, mesh]
    (polar components: F_R, F_phi, F_z, potential)."""
    mesh = nr * ngz * na
    if fld is None:
        fld = np.zeros((4, mesh), np.float64)
    _lib.orc3_fndf(nr, na, ngz, ng, np.ascontiguousarray(lg, np.int32),
                   np.ascontiguousarray(G, np.float64).reshape(-1),
                   np.ascontiguousarray(rho, np.float64), jrs, krs, jrf, krf, ntypes,
                   fld.reshape(-1))
    return fld


def direct_field3(nr, na, ngz, type_, D, rho):
    out = np.zeros(nr * ngz * na, np.float64)
    _lib.orc3_direct_field(nr, na, ngz, type_, np.ascontiguousarray(D),
                           np.ascontiguousarray(rho, np.float64), out)
    return out


def zmax(dzg, ngz, prec="d", jmass=2):
    """zm = dzg (ngz + 1 - jmass) / 2 (setgrd.f:146-150)."""
    if prec == "f":
        return float(np.float32(np.float32(0.5) * np.float32(dzg)) *
                     np.float32(ngz + 1 - jmass))
    return 0.5 * dzg * float(ngz + 1 - jmass)


def corners27(cell, na, ngz):
    """The 27 cells about each central cell (massgn.f:182-195), spoke offset
    fastest, then plane, then ring, as flat indices."""
    c = np.asarray(cell, np.int64)
    m0 = np.where(c % na == 0, c - 1 + na, c - 1)
    m2 = np.where((c + 1) % na == 0, c + 1 - na, c + 1)
    m = np.stack([m0, c, m2], 1)
    out = np.empty((len(c), 27), np.int64)
    for k in range(3):
        for j in range(3):
            for i in range(3):
                out[:, 9 * k + 3 * j + i] = m[:, i] + (j - 1) * na + (k - 1) * ngz * na
    return out


class P3D(_o.P2D):
    """Oracle state for one 3-D polar-grid model.  Shares the zone and time
    bookkeeping of the p2d oracle; mass arrays and fields are over
    mesh = nr * ngz * na."""

    def __init__(self, nr, na, ngz, dzg, hash_, softl, tsoft=1, nsect=1, skip=(),
                 nzones=1, nstep=(1,), zbr=(0.0,), pmass=1.0, offanal=True,
                 fixforce=None, prec="d", use_fx=True, qbits=None, green_tables=None,
                 jmass=2):
        self.nr, self.na, self.ngz, self.nsect = nr, na, ngz, nsect
        self.mesh = nr * ngz * na
        self.mmax = na // 2 + 1
        self.ng = min(hash_ + 1, self.mmax)
        self.prec = prec
        self.dt = np.float32 if prec == "f" else np.float64
        self.alpha = _o.alpha(na, prec)
        self.jmass = jmass
        self.nw = 27 if jmass == 3 else 8
        # rgrid = grofu(nr - jmass/2) (setgrd.f:135-136)
        self.rgrid = _o.grofu(self.dt(self.dt(nr) - self.dt(0.5) * self.dt(jmass)),
                              self.alpha, prec)
        self.thmax = self.dt(na) / self.dt(nsect)
        self.dzg = float(self.dt(dzg))
        self.zm = zmax(dzg, ngz, prec, jmass)
        self.softl, self.tsoft = softl, tsoft
        self.lg = filter_flags(na, self.ng, nsect, skip)
        self.nzones = nzones
        self.nstep = list(nstep)
        self.zbr = list(zbr)
        self.pmass = pmass
        self.offanal = offanal
        self.fixforce = fixforce
        self.use_fx = use_fx
        self.qbits = qbits
        if green_tables is None:
            G, _ = green3(nr, na, ngz, self.ng, self.lg, dzg, softl, tsoft)
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
        self.fld = np.zeros((4, self.mesh), np.float64)
        self.noff = 0

    def cfg(self):
        c = Cfg3()
        c.nr, c.na, c.ngz, c.nsect = self.nr, self.na, self.ngz, self.nsect
        c.nzones, c.mzone = self.nzones, self.mzone
        c.offanal = int(bool(self.offanal))
        c.jmass = self.jmass
        for i in range(self.nzones):
            c.nstep[i] = self.nstep[i]
            c.zbr[i] = self.zbr[i]
        c.alpha, c.rgrid, c.thmax = float(self.alpha), float(self.rgrid), float(self.thmax)
        c.zm, c.dzg = float(self.zm), float(self.dzg)
        if self.fixforce is not None:
            dr, val, der = self.fixforce
            self._fv = np.ascontiguousarray(val, np.float64)
            self._fd = np.ascontiguousarray(der, np.float64)
            c.fixforce, c.nfr, c.fr_dr = 1, len(self._fv), dr
            c.fr_val, c.fr_der = self._fv.ctypes.data, self._fd.ctypes.data
        return c

    def _fn(self, name):
        return getattr(_lib, name + ("_f" if self.prec == "f" else "_d"))

    # -- particles ---------------------------------------------------------
    def load(self, x, y, z, vx, vy, vz, iz=None):
        dt = self.dt
        self.x, self.y, self.z = (np.ascontiguousarray(a, dt).copy() for a in (x, y, z))
        self.vx, self.vy, self.vz = (np.ascontiguousarray(a, dt).copy() for a in (vx, vy, vz))
        n = len(self.x)
        self.iz = np.ascontiguousarray(np.ones(n, np.int32) if iz is None else iz,
                                       np.int32).copy()
        if self.qbits is None:
            self.qbits = _o.qbits_for(n)
        self.noff = int(np.sum(self.iz == self.nzones + 1))

    def weight(self, x, y, z, fold=0):
        dt = self.dt
        x, y, z = (np.ascontiguousarray(a, dt) for a in (x, y, z))
        n = len(x)
        cell = np.empty(n, np.int32)
        wt = np.empty((n, self.nw), dt)
        rr = np.empty(n, dt)
        c = self.cfg()
        self._fn("orc3_weight")(C.byref(c), n, x, y, z, int(fold), cell, wt, rr)
        return cell, wt, rr

    def getacc(self, field=None):
        """Accelerations (n, 3) of every particle from a Cartesian field
        (3, mesh) in the working precision (default: the solved one)."""
        if field is None:
            field = self.fld[:3].astype(self.dt)
        f = np.ascontiguousarray(field, self.dt).reshape(-1)
        acc = np.zeros((len(self.x), 3), self.dt)
        c = self.cfg()
        self._fn("orc3_getacc")(C.byref(c), len(self.x), self.x, self.y, self.z,
                                self.iz, f, acc)
        return acc

    def masset(self):
        self.mzone = self.nzones
        self._masscl(False)
        c = self.cfg()
        self._fn("orc3_massgn")(C.byref(c), len(self.x), self.x, self.y, self.z, self.iz,
                                self.grdmss.reshape(-1), self.mfx.reshape(-1),
                                self.qbits, self.jrad, self.krad)
        self._zoncmb_scaled()

    def findf(self, lpot=False):
        """findf -> p3fndf -> polcat (findf.f:40-60, p3fndf.f:37-104,
        polcat.f:22-40).  The solve replaces rings jrf..krf; polcat then
        turns every ring, as the reference's does, so on a partial step the
        rings outside the field range are turned once more (no moving
        particle reads them)."""
        jrs, krs = self.source_range()
        jrf, krf = self.field_range(lpot)
        ntypes = 4 if lpot else 3
        fndf3(self.nr, self.na, self.ngz, self.ng, self.lg, self.G, self.rho,
              jrs, krs, jrf, krf, ntypes, self.fld)
        _lib.orc_polcat(self.nr * self.ngz, self.na, self.fld.reshape(-1))
        self.ranges = (jrs, krs, jrf, krf)
        return self.fld

    def step(self, field=None):
        self._set_mzone()
        self._masscl(True)
        if field is None:
            field = self.fld[:3].astype(self.dt)
        f = np.ascontiguousarray(field, self.dt).reshape(-1)
        c = self.cfg()
        self.moved = np.zeros(len(self.x), np.uint8)
        dn = self._fn("orc3_push")(C.byref(c), len(self.x), self.x, self.y, self.z,
                                   self.vx, self.vy, self.vz, self.iz, f,
                                   self.grdmss.reshape(-1), self.mfx.reshape(-1),
                                   self.qbits, self.jrad, self.krad, self.moved)
        self.noff += dn
        self.istep += 1
        self._zoncmb_scaled()
        return dn

    def tmcent(self, start=True, field=None):
        if field is None:
            field = self.fld[:3].astype(self.dt)
        f = np.ascontiguousarray(field, self.dt).reshape(-1)
        c = self.cfg()
        self._fn("orc3_tmcent")(C.byref(c), len(self.x), self.x, self.y, self.z,
                                self.vx, self.vy, self.vz, self.iz, f, int(bool(start)))

    def integrals(self, fixpot=None, centred=False):
        """icheck's global integrals (SRC/lib15/icheck.f:24-70) in FP64 numpy,
        as galhip_integrals defines them for the 3-D grid: mass, m x position
        (3), momentum (3), sum m v^2, Lz, sum m Phi_grid (8-corner
        interpolation of the last potential with the particle's own weights),
        sum m Phi_fixed of the spherical radius.  centred: v + a/2 from the
        current field before the zone step factor."""
        x, y, z = (a.astype(np.float64) for a in (self.x, self.y, self.z))
        iz = self.iz
        zs = np.minimum(iz, self.nzones)
        sf = 1.0 / np.asarray(self.nstep, np.float64)[np.maximum(zs, 1) - 1]
        v = [self.vx, self.vy, self.vz]
        if centred:
            acc = self.getacc()
            v = [v[d] + self.dt(0.5) * acc[:, d] for d in range(3)]
        vx, vy, vz = (a.astype(np.float64) * sf for a in v)
        on = iz >= 1
        pm = float(self.pmass)
        out = np.zeros(11)
        out[0] = pm * on.sum()
        out[1], out[2], out[3] = pm * x[on].sum(), pm * y[on].sum(), pm * z[on].sum()
        out[4], out[5], out[6] = pm * vx[on].sum(), pm * vy[on].sum(), pm * vz[on].sum()
        out[7] = pm * (vx[on] ** 2 + vy[on] ** 2 + vz[on] ** 2).sum()
        out[8] = pm * (x[on] * vy[on] - y[on] * vx[on]).sum()
        grid = (iz >= 1) & (iz <= self.nzones)
        cell, w, _ = self.weight(self.x[grid], self.y[grid], self.z[grid])
        na, S = self.na, self.ngz * self.na
        if self.jmass == 3:
            ic = corners27(cell, na, self.ngz)
        else:
            wrap = (cell + 1) % na == 0
            c1 = np.where(wrap, cell + 1 - na, cell + 1)
            ic = np.stack([cell, c1, cell + na, c1 + na, cell + S, c1 + S, cell + S + na,
                           c1 + S + na], 1)
        pot = np.append(self.fld[3].astype(np.float64), 0.0)
        ic = np.where((ic >= 0) & (ic < self.mesh), ic, self.mesh)
        out[9] = pm * (w.astype(np.float64) * pot[ic]).sum()
        if fixpot is not None:
            dr, val, der = fixpot
            out[10] = pm * _o.fixforce(np.sqrt(x[on] ** 2 + y[on] ** 2 + z[on] ** 2),
                                       dr, val, der).sum()
        return out
