"""The Fortran interfaces in include/galhip_f.f90 against the C ABI.

A small Fortran program built with amdflang and linked against
libgalhip.so fills galhip_p2d_params (and galhip_p3d_params) field by field
with values that galhip_p2dset (galhip_p3dset) rejects before any HIP call,
and prints the messages: each
one names the field the C side read, so a wrong offset or kind in the
Fortran type shows up as a wrong message.  No GPU is used.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "galaxy_amd")
FLANG = shutil.which("amdflang") or (
    "/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)

PROGRAM = r"""
program bindcheck
  use galhip_f
  implicit none
  type(galhip_p2d_params) :: p
  type(c_ptr) :: ctx
  integer(c_int) :: rc
  print '(a,i0)', 'sizeof ', c_sizeof(p)
  call reset(p)
  p%precision = 7
  call try(p, 'precision')
  call reset(p)
  p%nr = 2
  p%na = 9
  call try(p, 'grid')
  call reset(p)
  p%nzones = 11
  call try(p, 'nzones')
  call reset(p)
  p%nsect = 0
  call try(p, 'nsect')
  call reset(p)
  p%softl = -1.0d0
  call try(p, 'softl')
  call reset(p)
  p%tsoft = 3
  call try(p, 'tsoft')
  call reset(p)
  p%nskip = 17
  call try(p, 'nskip')
  call p3checks()
contains
  subroutine p3checks()
    type(galhip_p3d_params) :: q
    print '(a,i0)', 'sizeof3 ', c_sizeof(q)
    call reset3(q)
    q%ngz = 2
    call try3(q, 'p3grid')
    call reset3(q)
    q%dzg = 0.0d0
    call try3(q, 'p3dzg')
    call reset3(q)
    q%softl = 0.1d0
    call try3(q, 'p3softl')
    call reset3(q)
    q%nzones = 0
    call try3(q, 'p3nzones')
    call reset3(q)
    q%tsoft = 0
    call try3(q, 'p3tsoft')
    call reset3(q)
    q%precision = 3
    call try3(q, 'p3precision')
    call reset3(q)
    q%jmass = 4
    call try3(q, 'p3jmass')
  end subroutine p3checks
  subroutine reset3(q)
    type(galhip_p3d_params), intent(out) :: q
    q%nr = 16
    q%na = 32
    q%ngz = 8
    q%hash = 4
    q%nsect = 1
    q%nskip = 0
    q%skip = 0
    q%dzg = 0.25d0
    q%softl = 0.5d0
    q%tsoft = 1
    q%nzones = 1
    q%nstep = 1
    q%zbr = 0.0d0
    q%offanal = 1
    q%pmass = 1.0d0
    q%precision = GALHIP_FP64
    q%jmass = 2
  end subroutine reset3
  subroutine try3(q, what)
    type(galhip_p3d_params), intent(in) :: q
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: i
    character(len=200) :: line
    rc = galhip_p3dset(q, ctx)
    call c_f_pointer(galhip_last_error(), msg, [200])
    line = ''
    do i = 1, 200
      if (msg(i) == c_null_char) exit
      line(i:i) = msg(i)
    end do
    print '(a,1x,i0,1x,a)', what, rc, trim(line)
  end subroutine try3
  subroutine reset(q)
    type(galhip_p2d_params), intent(out) :: q
    q%nr = 16
    q%na = 32
    q%hash = 4
    q%nsect = 1
    q%nskip = 0
    q%skip = 0
    q%softl = 0.5d0
    q%tsoft = 1
    q%nzones = 1
    q%nstep = 1
    q%zbr = 0.0d0
    q%offanal = 1
    q%pmass = 1.0d0
    q%precision = GALHIP_FP64
  end subroutine reset
  subroutine try(q, what)
    type(galhip_p2d_params), intent(in) :: q
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: i
    character(len=200) :: line
    rc = galhip_p2dset(q, ctx)
    call c_f_pointer(galhip_last_error(), msg, [200])
    line = ''
    do i = 1, 200
      if (msg(i) == c_null_char) exit
      line(i:i) = msg(i)
    end do
    print '(a,1x,i0,1x,a)', what, rc, trim(line)
  end subroutine try
end program bindcheck
"""


@pytest.mark.skipif(FLANG is None, reason="amdflang not installed")
def test_fortran_interfaces_match_c_abi(tmp_path):
    lib = os.path.join(LIBDIR, "libgalhip.so")
    if not os.path.exists(lib):
        pytest.skip("libgalhip.so not built")
    src = tmp_path / "bindcheck.f90"
    src.write_text(PROGRAM)
    exe = tmp_path / "bindcheck"
    subprocess.run([FLANG, "-J", str(tmp_path),
                    os.path.join(ROOT, "include", "galhip_f.f90"), str(src),
                    "-o", str(exe), "-L" + LIBDIR, "-lgalhip",
                    "-Wl,-rpath," + LIBDIR], check=True, cwd=tmp_path,
                   capture_output=True, timeout=300)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True,
                         timeout=120).stdout.splitlines()
    got = {ln.split()[0]: ln.split(None, 2)[1:] for ln in out}
    assert got["sizeof"] == ["248"]
    expect = {
        "precision": "precision must be",
        "grid": "grid 2 x 9 out of range",
        "nzones": "nzones 11 out of range",
        "nsect": "nsect not set",
        "softl": "Softening length",
        "tsoft": "Unrecognized softening rule",
        "nskip": "nskip out of range",
    }
    for k, msg in expect.items():
        rc, text = got[k]
        assert int(rc) == -1, (k, rc, text)
        assert text.startswith(msg), (k, text)
    # galhip_p3d_params: 6 + 16 ints, 2 doubles, 2 + 10 ints, 10 doubles,
    # 1 int (+ pad), 1 double, 1 int (+ pad) = 256 bytes
    assert got["sizeof3"] == ["256"]
    expect3 = {
        "p3grid": (-1, "grid 16 x 32 x 2 out of range"),
        "p3dzg": (-1, "z spacing must be positive"),
        "p3softl": (-3, "P3DSET: Softening length cannot be less than z-spacing"),
        "p3nzones": (-1, "nzones 0 out of range"),
        "p3tsoft": (-1, "Unrecognized softening rule"),
        "p3precision": (-1, "precision must be"),
        "p3jmass": (-1, "P3DSET: Interpolation scheme must be 2 or 3"),
    }
    for k, (code, msg) in expect3.items():
        rc, text = got[k]
        assert int(rc) == code, (k, rc, text)
        assert text.startswith(msg), (k, text)
