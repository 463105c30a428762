"""ctypes binding of libgalhip.so (declarations in include/galhip.h).

The library is built in-tree by galaxy_amd/build.py.  Loading fails loudly
when it is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libgalhip.so")

GALHIP_FP32 = 4
GALHIP_FP64 = 8
MAXZONES = 10
MAXSKIP = 16


class P2DParams(C.Structure):
    """galhip_p2d_params"""
    _fields_ = [
        ("nr", C.c_int32), ("na", C.c_int32), ("hash", C.c_int32),
        ("nsect", C.c_int32), ("nskip", C.c_int32),
        ("skip", C.c_int32 * MAXSKIP),
        ("softl", C.c_double), ("tsoft", C.c_int32),
        ("nzones", C.c_int32), ("nstep", C.c_int32 * MAXZONES),
        ("zbr", C.c_double * MAXZONES),
        ("offanal", C.c_int32), ("pmass", C.c_double),
        ("precision", C.c_int32),
    ]


class P3DParams(C.Structure):
    """galhip_p3d_params"""
    _fields_ = [
        ("nr", C.c_int32), ("na", C.c_int32), ("ngz", C.c_int32),
        ("hash", C.c_int32), ("nsect", C.c_int32), ("nskip", C.c_int32),
        ("skip", C.c_int32 * MAXSKIP),
        ("dzg", C.c_double), ("softl", C.c_double), ("tsoft", C.c_int32),
        ("nzones", C.c_int32), ("nstep", C.c_int32 * MAXZONES),
        ("zbr", C.c_double * MAXZONES),
        ("offanal", C.c_int32), ("pmass", C.c_double),
        ("precision", C.c_int32), ("jmass", C.c_int32),
    ]


_vp = C.c_void_p
_i32 = C.c_int32
_i64 = C.c_int64
_dbl = C.c_double

# name -> argtypes (all return int status, except galhip_last_error)
SIGNATURES = {
    "galhip_p2dset": [C.POINTER(P2DParams), C.POINTER(_vp)],
    "galhip_free": [_vp],
    "galhip_geometry": [_vp, _vp],
    "galhip_greenm": [_vp],
    "galhip_get_green": [_vp, _vp],
    "galhip_set_green": [_vp, _vp],
    "galhip_set_fixforce": [_vp, _i32, _dbl, _vp, _vp],
    "galhip_restrt": [_vp, _i64, _vp, _vp, _vp, _vp, _vp],
    "galhip_unload": [_vp, _vp, _vp, _vp, _vp, _vp],
    "galhip_set_step": [_vp, _i32, _vp],
    "galhip_get_step": [_vp, _vp, _vp, _vp],
    "galhip_masset": [_vp],
    "galhip_findf": [_vp, _i32],
    "galhip_p2fndf": [_vp, _i32, _i32, _i32],
    "galhip_step": [_vp],
    "galhip_run": [_vp, _i32],
    "galhip_tmcent": [_vp, _i32],
    "galhip_weight": [_vp, _i64, _vp, _vp, _i32, _vp, _vp],
    "galhip_grdacc": [_vp, _i64, _vp, _vp, _vp],
    "galhip_get_mass": [_vp, _vp, _vp, _vp, _vp, _vp],
    "galhip_set_mass": [_vp, _vp, _vp, _vp, _vp],
    "galhip_get_rho": [_vp, _vp],
    "galhip_get_field": [_vp, _vp],
    "galhip_set_field": [_vp, _vp],
    "galhip_get_ranges": [_vp, _vp],
    "galhip_get_noff": [_vp, _vp],
    "galhip_set_fixpot": [_vp, _i32, _dbl, _vp, _vp],
    "galhip_integrals": [_vp, _i32, _vp],
    "galhip_integrals_noalloc": [_vp, _i32, _vp],
    "galhip_lgspia": [_vp, _i32, _i32, _dbl, _vp],
    "galhip_set_sort_interval": [_vp, _i32],
    "galhip_set_pmass": [_vp, _dbl],
    "galhip_sort": [_vp],
    "galhip_get_order": [_vp, _vp],
    "galhip_set_profile": [_vp, _i32],
    "galhip_get_profile": [_vp, _vp, _vp, _vp],
    "galhip_sync": [_vp],
    "galhip_p3dset": [C.POINTER(P3DParams), C.POINTER(_vp)],
    "galhip_restrt3": [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "galhip_unload3": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "galhip_weight3": [_vp, _i64, _vp, _vp, _vp, _i32, _vp, _vp],
    "galhip_grdacc3": [_vp, _i64, _vp, _vp, _vp, _vp],
    "galhip_getacc": [_vp, _i64, _vp, _vp, _vp, _vp],
    "galhip_stpgrp": [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp],
    "galhip_massgn": [_vp, _i64, _vp, _vp, _vp],
    "galhip_ensset": [C.POINTER(P2DParams), _i32, _vp, C.POINTER(_vp)],
    "galhip_ens_restrt": [_vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "galhip_ens_unload": [_vp, _vp, _vp, _vp, _vp, _vp],
    "galhip_ens_get_mass": [_vp, _i32, _vp, _vp, _vp, _vp, _vp],
    "galhip_ens_get_field": [_vp, _i32, _vp],
    "galhip_ens_set_field": [_vp, _i32, _vp],
    "galhip_ens_get_ranges": [_vp, _i32, _vp],
    "galhip_ens_get_noff": [_vp, _vp],
    "galhip_ens_integrals": [_vp, _i32, _vp],
    "galhip_set_graph": [_vp, _i32],
    "galhip_set_contraction": [_vp, _i32],
    "galhip_set_compensated": [_vp, _i32],
    "galhip_get_position_lo": [_vp, _vp, _vp],
    "galhip_get_contraction": [_vp, _vp, _vp, _vp],
    "galhip_adj_record": [_vp, _i32],
    "galhip_adj_restore": [_vp, _i32],
    "galhip_adj_set": [_vp, _vp],
    "galhip_adj_get": [_vp, _vp],
    "galhip_set_fixforce_dpar": [_vp, _i32, _dbl, _vp, _vp],
    "galhip_adj_step": [_vp, _vp],
    "galhip_adj_get_par": [_vp, _vp],
    "galhip_adj_seed_a2": [_vp, _dbl, _dbl, _dbl, _vp],
    "galhip_adj_field": [_vp, _vp, _vp],
    "galhip_adj_get_profile": [_vp, _vp, _vp],
    "galhip_tl_set": [_vp, _vp],
    "galhip_tl_get": [_vp, _vp],
    "galhip_tl_step": [_vp],
    "galhip_tl_field": [_vp, _vp, _vp],
}

_lib = None


def load(path: str = None) -> C.CDLL:
    """Load libgalhip.so once; raises if it has not been built.  The
    environment variable GALHIP_LIBRARY selects another build of the same
    library (for launch-shape tuning)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.environ.get("GALHIP_LIBRARY") or LIBPATH
    if not os.path.exists(path):
        raise RuntimeError(
            "libgalhip.so not found at %s: build it with "
            "`python galaxy_amd/build.py` (hipcc, gfx950)" % path)
    lib = C.CDLL(path)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.galhip_last_error.argtypes = []
    lib.galhip_last_error.restype = C.c_char_p
    _lib = lib
    return lib


class GalhipError(RuntimeError):
    """A non-zero status from libgalhip (the reference would crash())."""


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().galhip_last_error().decode(errors="replace")
        raise GalhipError("%s failed (%d): %s" % (what, rc, msg))
