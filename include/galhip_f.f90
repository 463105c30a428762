! galhip_f.f90 -- Fortran 2003 (ISO_C_BINDING) interfaces to libgalhip.so,
! one for each entry point of galhip.h, for GALAXY's Fortran host code.
!
! The reference swaps implementations of its subroutines by link order
! (SRC/progs/Makefile:22,25,55 put lib15_mpi ahead of main15.a).  A shim that
! replaces findf, step and masset (SRC/lib15/findf.f:1, step.f:1,
! masset.f:1) uses this module and links ahead of main15.a with
! -L<dir> -lgalhip.  Floating point arguments are passed as type(c_ptr)
! (c_loc of a real or real*8 array), because their kind follows the
! context's precision.
module galhip_f
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: GALHIP_FP32 = 4, GALHIP_FP64 = 8
  integer(c_int32_t), parameter :: GALHIP_MAXZONES = 10, GALHIP_MAXSKIP = 16
  integer(c_int), parameter :: GALHIP_OK = 0, GALHIP_EINVAL = -1, &
       GALHIP_EHIP = -2, GALHIP_ERANGE = -3

  type, bind(c) :: galhip_p2d_params
     integer(c_int32_t) :: nr, na, hash, nsect, nskip
     integer(c_int32_t) :: skip(GALHIP_MAXSKIP)
     real(c_double) :: softl
     integer(c_int32_t) :: tsoft, nzones
     integer(c_int32_t) :: nstep(GALHIP_MAXZONES)
     real(c_double) :: zbr(GALHIP_MAXZONES)
     integer(c_int32_t) :: offanal
     real(c_double) :: pmass
     integer(c_int32_t) :: precision
  end type galhip_p2d_params

  type, bind(c) :: galhip_p3d_params
     integer(c_int32_t) :: nr, na, ngz, hash, nsect, nskip
     integer(c_int32_t) :: skip(GALHIP_MAXSKIP)
     real(c_double) :: dzg, softl
     integer(c_int32_t) :: tsoft, nzones
     integer(c_int32_t) :: nstep(GALHIP_MAXZONES)
     real(c_double) :: zbr(GALHIP_MAXZONES)
     integer(c_int32_t) :: offanal
     real(c_double) :: pmass
     integer(c_int32_t) :: precision, jmass
  end type galhip_p3d_params

  interface
     integer(c_int) function galhip_p2dset(p, ctx) bind(c, name='galhip_p2dset')
       import :: c_int, c_ptr, galhip_p2d_params
       type(galhip_p2d_params), intent(in) :: p
       type(c_ptr), intent(out) :: ctx
     end function galhip_p2dset
     integer(c_int) function galhip_free(ctx) bind(c, name='galhip_free')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function galhip_free
     type(c_ptr) function galhip_last_error() bind(c, name='galhip_last_error')
       import :: c_ptr
     end function galhip_last_error
     integer(c_int) function galhip_geometry(ctx, out) bind(c, name='galhip_geometry')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: out(*)
     end function galhip_geometry
     integer(c_int) function galhip_greenm(ctx) bind(c, name='galhip_greenm')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function galhip_greenm
     integer(c_int) function galhip_get_green(ctx, g) bind(c, name='galhip_get_green')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: g(*)
     end function galhip_get_green
     integer(c_int) function galhip_set_green(ctx, g) bind(c, name='galhip_set_green')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: g(*)
     end function galhip_set_green
     integer(c_int) function galhip_set_fixforce(ctx, n, dr, val, der) &
          bind(c, name='galhip_set_fixforce')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n
       real(c_double), value :: dr
       real(c_double), intent(in) :: val(*), der(*)
     end function galhip_set_fixforce
     integer(c_int) function galhip_restrt(ctx, n, x, y, vx, vy, iz) &
          bind(c, name='galhip_restrt')
       import :: c_int, c_ptr, c_int64_t, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: n
       type(c_ptr), value :: x, y, vx, vy
       integer(c_int32_t), intent(in) :: iz(*)
     end function galhip_restrt
     integer(c_int) function galhip_unload(ctx, x, y, vx, vy, iz) &
          bind(c, name='galhip_unload')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       type(c_ptr), value :: x, y, vx, vy, iz
     end function galhip_unload
     integer(c_int) function galhip_set_step(ctx, istep, lstep) &
          bind(c, name='galhip_set_step')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: istep
       type(c_ptr), value :: lstep
     end function galhip_set_step
     integer(c_int) function galhip_get_step(ctx, istep, mzone, lstep) &
          bind(c, name='galhip_get_step')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: istep, mzone
       type(c_ptr), value :: lstep
     end function galhip_get_step
     integer(c_int) function galhip_masset(ctx) bind(c, name='galhip_masset')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function galhip_masset
     integer(c_int) function galhip_findf(ctx, lpot) bind(c, name='galhip_findf')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: lpot
     end function galhip_findf
     integer(c_int) function galhip_p2fndf(ctx, jzone, kzone, lpot) &
          bind(c, name='galhip_p2fndf')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: jzone, kzone, lpot
     end function galhip_p2fndf
     integer(c_int) function galhip_step(ctx) bind(c, name='galhip_step')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function galhip_step
     integer(c_int) function galhip_tmcent(ctx, start) bind(c, name='galhip_tmcent')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: start
     end function galhip_tmcent
     integer(c_int) function galhip_run(ctx, nsteps) bind(c, name='galhip_run')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nsteps
     end function galhip_run
     integer(c_int) function galhip_weight(ctx, n, x, y, old, cell, wt) &
          bind(c, name='galhip_weight')
       import :: c_int, c_ptr, c_int64_t, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: n
       type(c_ptr), value :: x, y
       integer(c_int32_t), value :: old
       integer(c_int32_t), intent(out) :: cell(*)
       type(c_ptr), value :: wt
     end function galhip_weight
     integer(c_int) function galhip_grdacc(ctx, n, x, y, acc) &
          bind(c, name='galhip_grdacc')
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: n
       type(c_ptr), value :: x, y, acc
     end function galhip_grdacc
     integer(c_int) function galhip_getacc(ctx, n, x, y, iz, acc) &
          bind(c, name='galhip_getacc')
       import :: c_int, c_ptr, c_int64_t, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: n
       type(c_ptr), value :: x, y, acc
       integer(c_int32_t), intent(in) :: iz(*)
     end function galhip_getacc
     integer(c_int) function galhip_stpgrp(ctx, n, x, y, vx, vy, acc, iz) &
          bind(c, name='galhip_stpgrp')
       import :: c_int, c_ptr, c_int64_t, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: n
       type(c_ptr), value :: x, y, vx, vy, acc
       integer(c_int32_t), intent(inout) :: iz(*)
     end function galhip_stpgrp
     integer(c_int) function galhip_massgn(ctx, n, x, y, iz) &
          bind(c, name='galhip_massgn')
       import :: c_int, c_ptr, c_int64_t, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: n
       type(c_ptr), value :: x, y
       integer(c_int32_t), intent(in) :: iz(*)
     end function galhip_massgn
     integer(c_int) function galhip_get_mass(ctx, mfx, mozfx, jrad, krad, qbits) &
          bind(c, name='galhip_get_mass')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       type(c_ptr), value :: mfx, mozfx
       integer(c_int32_t), intent(out) :: jrad(*), krad(*), qbits
     end function galhip_get_mass
     integer(c_int) function galhip_set_mass(ctx, mfx, mozfx, jrad, krad) &
          bind(c, name='galhip_set_mass')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       type(c_ptr), value :: mfx, mozfx
       integer(c_int32_t), intent(in) :: jrad(*), krad(*)
     end function galhip_set_mass
     integer(c_int) function galhip_get_rho(ctx, rho) bind(c, name='galhip_get_rho')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, rho
     end function galhip_get_rho
     integer(c_int) function galhip_get_field(ctx, fld) bind(c, name='galhip_get_field')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, fld
     end function galhip_get_field
     integer(c_int) function galhip_set_field(ctx, fld) bind(c, name='galhip_set_field')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, fld
     end function galhip_set_field
     integer(c_int) function galhip_get_ranges(ctx, r4) bind(c, name='galhip_get_ranges')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: r4(4)
     end function galhip_get_ranges
     integer(c_int) function galhip_get_noff(ctx, noff) bind(c, name='galhip_get_noff')
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), intent(out) :: noff
     end function galhip_get_noff
     integer(c_int) function galhip_set_fixpot(ctx, n, dr, val, der) &
          bind(c, name='galhip_set_fixpot')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n
       real(c_double), value :: dr
       real(c_double), intent(in) :: val(*), der(*)
     end function galhip_set_fixpot
     integer(c_int) function galhip_integrals(ctx, centred, out) &
          bind(c, name='galhip_integrals')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: centred
       real(c_double), intent(out) :: out(9)
     end function galhip_integrals
     integer(c_int) function galhip_integrals_noalloc(ctx, centred, out) &
          bind(c, name='galhip_integrals_noalloc')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: centred
       real(c_double), intent(out) :: out(*)
     end function galhip_integrals_noalloc
     integer(c_int) function galhip_lgspia(ctx, np, nm, lscale, out) &
          bind(c, name='galhip_lgspia')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: np, nm
       real(c_double), value :: lscale
       real(c_double), intent(out) :: out(*)
     end function galhip_lgspia
     integer(c_int) function galhip_set_sort_interval(ctx, k) &
          bind(c, name='galhip_set_sort_interval')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: k
     end function galhip_set_sort_interval
     integer(c_int) function galhip_set_pmass(ctx, pmass) bind(c, name='galhip_set_pmass')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: pmass
     end function galhip_set_pmass
     integer(c_int) function galhip_sort(ctx) bind(c, name='galhip_sort')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function galhip_sort
     integer(c_int) function galhip_get_order(ctx, ids) bind(c, name='galhip_get_order')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: ids(*)
     end function galhip_get_order
     integer(c_int) function galhip_set_profile(ctx, on) bind(c, name='galhip_set_profile')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: on
     end function galhip_set_profile
     integer(c_int) function galhip_get_profile(ctx, push_ms, launches, pushes) &
          bind(c, name='galhip_get_profile')
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: push_ms
       integer(c_int64_t), intent(out) :: launches, pushes
     end function galhip_get_profile
     integer(c_int) function galhip_sync(ctx) bind(c, name='galhip_sync')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function galhip_sync
     integer(c_int) function galhip_p3dset(p, ctx) bind(c, name='galhip_p3dset')
       import :: c_int, c_ptr, galhip_p3d_params
       type(galhip_p3d_params), intent(in) :: p
       type(c_ptr), intent(out) :: ctx
     end function galhip_p3dset
     integer(c_int) function galhip_restrt3(ctx, n, x, y, z, vx, vy, vz, iz) &
          bind(c, name='galhip_restrt3')
       import :: c_int, c_ptr, c_int64_t, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: n
       type(c_ptr), value :: x, y, z, vx, vy, vz
       integer(c_int32_t), intent(in) :: iz(*)
     end function galhip_restrt3
     integer(c_int) function galhip_unload3(ctx, x, y, z, vx, vy, vz, iz) &
          bind(c, name='galhip_unload3')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       type(c_ptr), value :: x, y, z, vx, vy, vz, iz
     end function galhip_unload3
     integer(c_int) function galhip_weight3(ctx, n, x, y, z, old, cell, wt) &
          bind(c, name='galhip_weight3')
       import :: c_int, c_ptr, c_int64_t, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: n
       type(c_ptr), value :: x, y, z
       integer(c_int32_t), value :: old
       integer(c_int32_t), intent(out) :: cell(*)
       type(c_ptr), value :: wt
     end function galhip_weight3
     integer(c_int) function galhip_grdacc3(ctx, n, x, y, z, acc) &
          bind(c, name='galhip_grdacc3')
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: n
       type(c_ptr), value :: x, y, z, acc
     end function galhip_grdacc3
     ! ensembles of p2d discs (no counterpart in the reference)
     integer(c_int) function galhip_ensset(p, nmem, pmass, ctx) bind(c, name='galhip_ensset')
       import :: c_int, c_ptr, c_int32_t, c_double, galhip_p2d_params
       type(galhip_p2d_params), intent(in) :: p
       integer(c_int32_t), value :: nmem
       real(c_double), intent(in) :: pmass(*)
       type(c_ptr), intent(out) :: ctx
     end function galhip_ensset
     integer(c_int) function galhip_ens_restrt(ctx, count, x, y, vx, vy, iz) &
          bind(c, name='galhip_ens_restrt')
       import :: c_int, c_ptr, c_int64_t, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), intent(in) :: count(*)
       type(c_ptr), value :: x, y, vx, vy
       integer(c_int32_t), intent(in) :: iz(*)
     end function galhip_ens_restrt
     integer(c_int) function galhip_ens_unload(ctx, x, y, vx, vy, iz) &
          bind(c, name='galhip_ens_unload')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       type(c_ptr), value :: x, y, vx, vy, iz
     end function galhip_ens_unload
     integer(c_int) function galhip_ens_get_mass(ctx, member, mfx, mozfx, jrad, krad, qbits) &
          bind(c, name='galhip_ens_get_mass')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: member
       type(c_ptr), value :: mfx, mozfx, jrad, krad, qbits
     end function galhip_ens_get_mass
     integer(c_int) function galhip_ens_get_field(ctx, member, fld) &
          bind(c, name='galhip_ens_get_field')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: member
       type(c_ptr), value :: fld
     end function galhip_ens_get_field
     integer(c_int) function galhip_ens_set_field(ctx, member, fld) &
          bind(c, name='galhip_ens_set_field')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: member
       type(c_ptr), value :: fld
     end function galhip_ens_set_field
     integer(c_int) function galhip_ens_get_ranges(ctx, member, r4) &
          bind(c, name='galhip_ens_get_ranges')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: member
       integer(c_int32_t), intent(out) :: r4(4)
     end function galhip_ens_get_ranges
     integer(c_int) function galhip_ens_get_noff(ctx, noff) bind(c, name='galhip_ens_get_noff')
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), intent(out) :: noff(*)
     end function galhip_ens_get_noff
     integer(c_int) function galhip_ens_integrals(ctx, centred, out) &
          bind(c, name='galhip_ens_integrals')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: centred
       real(c_double), intent(out) :: out(*)
     end function galhip_ens_integrals
     ! execution and precision controls
     integer(c_int) function galhip_set_graph(ctx, on) bind(c, name='galhip_set_graph')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: on
     end function galhip_set_graph
     integer(c_int) function galhip_set_contraction(ctx, mode) &
          bind(c, name='galhip_set_contraction')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: mode
     end function galhip_set_contraction
     integer(c_int) function galhip_get_contraction(ctx, mode, variant, us) &
          bind(c, name='galhip_get_contraction')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: mode, variant
       real(c_double), intent(out) :: us(3)
     end function galhip_get_contraction
     integer(c_int) function galhip_set_compensated(ctx, on) &
          bind(c, name='galhip_set_compensated')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: on
     end function galhip_set_compensated
     integer(c_int) function galhip_get_position_lo(ctx, xlo, ylo) &
          bind(c, name='galhip_get_position_lo')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       type(c_ptr), value :: xlo, ylo
     end function galhip_get_position_lo
     ! adjoint of the p2d step (no counterpart in the reference)
     integer(c_int) function galhip_adj_record(ctx, slot) bind(c, name='galhip_adj_record')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: slot
     end function galhip_adj_record
     integer(c_int) function galhip_adj_restore(ctx, slot) bind(c, name='galhip_adj_restore')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: slot
     end function galhip_adj_restore
     integer(c_int) function galhip_adj_set(ctx, lam) bind(c, name='galhip_adj_set')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: lam(*)
     end function galhip_adj_set
     integer(c_int) function galhip_adj_get(ctx, lam) bind(c, name='galhip_adj_get')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: lam(*)
     end function galhip_adj_get
     integer(c_int) function galhip_set_fixforce_dpar(ctx, n, dr, val, der) &
          bind(c, name='galhip_set_fixforce_dpar')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n
       real(c_double), value :: dr
       real(c_double), intent(in) :: val(*), der(*)
     end function galhip_set_fixforce_dpar
     integer(c_int) function galhip_adj_step(ctx, gpar) bind(c, name='galhip_adj_step')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(inout) :: gpar(2)
     end function galhip_adj_step
     integer(c_int) function galhip_adj_get_par(ctx, gpar) bind(c, name='galhip_adj_get_par')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: gpar(2)
     end function galhip_adj_get_par
     integer(c_int) function galhip_adj_seed_a2(ctx, lscale, r1, r2, out) &
          bind(c, name='galhip_adj_seed_a2')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: lscale, r1, r2
       real(c_double), intent(out) :: out(4)
     end function galhip_adj_seed_a2
     integer(c_int) function galhip_adj_field(ctx, g, rbar) bind(c, name='galhip_adj_field')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: g(*)
       real(c_double), intent(out) :: rbar(*)
     end function galhip_adj_field
     integer(c_int) function galhip_adj_get_profile(ctx, ms, steps) &
          bind(c, name='galhip_adj_get_profile')
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: ms(2)
       integer(c_int64_t), intent(out) :: steps
     end function galhip_adj_get_profile
     integer(c_int) function galhip_tl_set(ctx, d) bind(c, name='galhip_tl_set')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: d(*)
     end function galhip_tl_set
     integer(c_int) function galhip_tl_get(ctx, d) bind(c, name='galhip_tl_get')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: d(*)
     end function galhip_tl_get
     integer(c_int) function galhip_tl_step(ctx) bind(c, name='galhip_tl_step')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function galhip_tl_step
     integer(c_int) function galhip_tl_field(ctx, rho, f) bind(c, name='galhip_tl_field')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: rho(*)
       real(c_double), intent(out) :: f(*)
     end function galhip_tl_field
  end interface
end module galhip_f
