"""The lib15-named override layer (galaxy_amd/fortran/galhip_shim.f).

The shim is compiled with amdflang against the reference's own include
files and aarrays module where they lie (SRC/lib15/inc, SRC/lib15/aarrays.f;
the objects go to a temporary directory, nothing is copied or kept), so
the COMMON blocks and module arrays it reads are the reference's.  Then:

  - its object defines findf_, step_, masset_, greenm_, p2fndf_, p3fndf_,
    polcat_ and tmcent_, the external names of the routines it replaces
    (SRC/lib15/findf.f:1, step.f:1, masset.f:1, greenm.f:1, p2fndf.f:1,
    p3fndf.f:1, polcat.f:1, tmcent.f:1);
  - linked ahead of a stand-in for main15.a (an archive with one member
    per routine, as the reference's library is built, carrying the same
    names and printing a marker if it were called), a small program that
    calls greenm resolves to the shim: greenm_ creates the context through
    galhip_p2dset with the grid set in the reference's COMMON blocks,
    which the library rejects before any HIP call, and the shim stops
    through crash( 'P2DSET', message ) with the library's message -- the
    reference's error convention (SRC/lib15/crash.f:23-37).

No GPU is used.  Skipped where the reference tree or amdflang is absent.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "galaxy_amd")
SHIM = os.path.join(ROOT, "galaxy_amd", "fortran", "galhip_shim.f")
LIB15 = "/root/reference/SRC/lib15"
FLANG = shutil.which("amdflang") or (
    "/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)

OVERRIDDEN = ["findf_", "step_", "masset_", "greenm_", "p2fndf_", "p3fndf_", "polcat_",
              "tmcent_"]

# main15.a stand-in: one member per routine, as the reference's library
# has one object per source file
STUBS = {
    "findf": "      subroutine findf( lpot )\n      logical lpot\n",
    "step": "      subroutine step\n",
    "masset": "      subroutine masset\n",
    "greenm": "      subroutine greenm( direct )\n      logical direct\n",
    "p2fndf": "      subroutine p2fndf( j, k )\n      integer j, k\n",
    "p3fndf": "      subroutine p3fndf( j, k )\n      integer j, k\n",
    "polcat": "      subroutine polcat\n",
    "tmcent": "      subroutine tmcent( start )\n      logical start\n",
}

OTHER_STUBS = {
    "crash": ("      subroutine crash( routine, reason )\n"
              "      character*(*) routine, reason\n"
              "      print '( a, 1x, a, 1x, a )', 'CRASH', routine, reason\n"
              "      stop\n      end\n"),
    "switch": "      subroutine switch( j )\n      integer j\n      end\n",
    "halfrc": "      real function halfrc( r )\n      real r\n      halfrc = 0\n      end\n",
    "axifrc": "      real function axifrc( r )\n      real r\n      axifrc = 0\n      end\n",
    "frcorr": "      real function frcorr( r )\n      real r\n      frcorr = 0\n      end\n",
}

MAIN = """      program shimcheck
      implicit none
      include 'inc/params.f'
      include 'inc/admin.f'
      include 'inc/grids.f'
      ngrid = 1
      p2d = .true.
      lfrv = .true.
      nr( 1 ) = 2
      na = 9
      ng = 1
      nsect = 1
      nzones = 1
      nstep( 1 ) = 1
      softl = 1.0
      tsoft = 1
      pmass = 1.0
      call greenm( .false. )
      print '( a )', 'NOT STOPPED'
      end
"""


def _run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (cmd, r.stdout, r.stderr)
    return r.stdout


@pytest.mark.skipif(FLANG is None, reason="amdflang not installed")
@pytest.mark.skipif(not os.path.isdir(os.path.join(LIB15, "inc")),
                    reason="reference tree not present")
def test_shim_compiles_and_overrides(tmp_path):
    lib = os.path.join(LIBDIR, "libgalhip.so")
    if not os.path.exists(lib):
        pytest.skip("libgalhip.so not built")
    t = str(tmp_path)
    # the reference's module, compiled where it lies (objects to tmp only)
    _run([FLANG, "-c", "-J", t, os.path.join(LIB15, "aarrays.f"), "-o", "aarrays.o"], t)
    _run([FLANG, "-c", "-J", t, os.path.join(ROOT, "include", "galhip_f.f90"),
          "-o", "galhip_f.o"], t)
    _run([FLANG, "-c", "-J", t, "-I", LIB15, SHIM, "-o", "galhip_shim.o"], t)
    syms = _run(["nm", "galhip_shim.o"], t).split("\n")
    defined = {ln.split()[-1] for ln in syms if " T " in ln}
    for s in OVERRIDDEN:
        assert s in defined, s
    # main15.a stand-in, one member per routine
    members = []
    for name, head in STUBS.items():
        src = tmp_path / ("%s.f" % name)
        src.write_text(head + "      print '( a )', 'REFERENCE %s'\n      end\n" % name)
        _run([FLANG, "-c", str(src), "-o", "%s.o" % name], t)
        members.append("%s.o" % name)
    for name, text in OTHER_STUBS.items():
        src = tmp_path / ("%s.f" % name)
        src.write_text(text)
        _run([FLANG, "-c", str(src), "-o", "%s.o" % name], t)
        members.append("%s.o" % name)
    _run(["ar", "rcs", "libmain15.a"] + members, t)
    (tmp_path / "main.f").write_text(MAIN)
    _run([FLANG, "-c", "-I", LIB15, "main.f", "-o", "main.o"], t)
    # the override layer first, then the library, then main15.a
    _run([FLANG, "main.o", "galhip_shim.o", "galhip_f.o", "aarrays.o",
          "-L" + LIBDIR, "-lgalhip", "-Wl,-rpath," + LIBDIR, "libmain15.a",
          "-o", "shimcheck"], t)
    exe_syms = _run(["nm", "shimcheck"], t)
    assert "galhip_open_" in exe_syms
    out = _run(["./shimcheck"], t)
    assert "REFERENCE" not in out, out
    line = [ln for ln in out.splitlines() if ln.startswith("CRASH")]
    assert line, out
    assert line[0].split(None, 2)[1] == "P2DSET", line
    assert "grid 2 x 9 out of range" in line[0], line
    assert "NOT STOPPED" not in out
