"""The C-ABI library loads without a GPU and exports every entry point
declared in include/galhip.h; the Python binding declares the same set.
No compute call is made here."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "galhip.h")
LIB = os.path.join(ROOT, "galaxy_amd", "libgalhip.so")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(galhip_\w+)\s*\(",
                                 text, re.M)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("galhip_p2dset", "galhip_greenm", "galhip_restrt",
                 "galhip_masset", "galhip_findf", "galhip_step", "galhip_run",
                 "galhip_weight", "galhip_grdacc", "galhip_last_error"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libgalhip.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_matches_header():
    from galaxy_amd import _lib
    names = set(declared()) - {"galhip_last_error"}
    assert names == set(_lib.SIGNATURES), names ^ set(_lib.SIGNATURES)


def test_params_struct_layout():
    """galhip_p2d_params as ctypes lays it out (natural C alignment)."""
    from galaxy_amd._lib import P2DParams
    off = {f[0]: getattr(P2DParams, f[0]).offset for f in P2DParams._fields_}
    assert off["skip"] == 20 and off["softl"] == 88
    assert off["zbr"] == 144 and ctypes.sizeof(P2DParams) == 248


@pytest.mark.skipif(not os.path.exists(LIB), reason="libgalhip.so not built")
def test_p2dset_rejects_grids_coarse_against_softening():
    """setgrd.f:181-199: the reference stops when the outermost cell is more
    than 5 softening lengths wide; galhip_p2dset returns GALHIP_ERANGE there
    (before any device call, so this runs without a GPU).  Widths here: the
    last ring of a 24 x 32 grid is rgrid - grofu(22) = 2.0 grid units."""
    from galaxy_amd import _lib
    lib = _lib.load()
    p = _lib.P2DParams()
    p.nr, p.na, p.hash, p.nsect, p.nskip = 24, 32, 2, 1, 0
    p.tsoft, p.nzones, p.offanal, p.pmass = 1, 1, 1, 1.0
    p.nstep[0] = 1
    p.precision = _lib.GALHIP_FP64
    import math
    a = 2 * math.pi / 32
    width = (math.exp(a * 23) - 1) - (math.exp(a * 22) - 1)
    ctx = ctypes.c_void_p()
    p.softl = width / 5.5
    rc = lib.galhip_p2dset(ctypes.byref(p), ctypes.byref(ctx))
    assert rc == -3 and b"SETGRD" in lib.galhip_last_error()


@pytest.mark.skipif(not os.path.exists(LIB), reason="libgalhip.so not built")
def test_p2dset_rejects_grids_out_of_range():
    """Grids outside 3 <= nr <= 4096, 4 <= na <= 2048 return GALHIP_EINVAL
    before any device call (nr <= 4096: the push kernel packs a deposit's
    zone, ring and spoke into one 32-bit word)."""
    from galaxy_amd import _lib
    lib = _lib.load()
    for nr, na in ((4097, 64), (2, 64), (64, 2050)):
        p = _lib.P2DParams()
        p.nr, p.na, p.hash, p.nsect, p.nskip = nr, na, 2, 1, 0
        p.tsoft, p.nzones, p.offanal, p.pmass, p.softl = 1, 1, 1, 1.0, 0.1
        p.nstep[0] = 1
        p.precision = _lib.GALHIP_FP32
        ctx = ctypes.c_void_p()
        rc = lib.galhip_p2dset(ctypes.byref(p), ctypes.byref(ctx))
        assert rc == -1 and b"out of range" in lib.galhip_last_error(), (nr, na, rc)


def test_fortran_module_declares_every_entry_point():
    """include/galhip_f.f90 binds every function galhip.h declares."""
    text = open(os.path.join(ROOT, "include", "galhip_f.f90")).read()
    bound = set(re.findall(r"bind\(c,\s*name='(galhip_\w+)'\)", text))
    missing = set(declared()) - bound - {"galhip_last_error"}
    assert not missing, sorted(missing)


DRIVER = os.path.join(ROOT, "galaxy_amd", "galhip_galaxy")


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")
def test_driver_rejects_bad_input(tmp_path):
    """The native driver's argument and file checks (before any GPU call)."""
    import subprocess
    r = subprocess.run([DRIVER], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([DRIVER, str(tmp_path / "none.pcs"), str(tmp_path / "o.pcs"),
                        "nr=8", "na=16", "hash=2", "soft=1", "lscale=1", "ts=0.1", "steps=1"],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "cannot read" in r.stderr
    r = subprocess.run([DRIVER, "a", "b", "steps=1", "bogus=3"], capture_output=True, text=True)
    assert r.returncode == 2 and "unknown key" in r.stderr
