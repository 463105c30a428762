c galhip_shim.f -- lib15-named overrides that run GALAXY's p2d and p3d
c   polar-grid step on the GPU through libgalhip.so (include/galhip.h,
c   Fortran interfaces in include/galhip_f.f90).
c
c GALAXY swaps implementations of its routines by link order: an archive
c   named ahead of main15.a on the link line supplies routines of the
c   same name (SRC/progs/Makefile:22,25,55 does this for lib15_mpi).
c   Build this file against the reference's own includes and aarrays
c   module and link it first:
c     flang -c -I<lib15> -J<mod> galhip_f.f90 galhip_shim.f
c     ... galhip_shim.o galhip_f.o -L<dir> -lgalhip main15.a ...
c
c The overrides keep the reference's argument lists:
c   greenm( direct )        SRC/lib15/greenm.f:1   context + tables
c   masset                  SRC/lib15/masset.f:1
c   findf( lpot )           SRC/lib15/findf.f:1    solve + polcat
c   p2fndf( jzone, kzone )  SRC/lib15/p2fndf.f:1   solve + polcat
c   p3fndf( jzone, kzone )  SRC/lib15/p3fndf.f:1   solve + polcat
c   polcat                  SRC/lib15/polcat.f:1   done by the solve
c   step                    SRC/lib15/step.f:1
c   tmcent( start )         SRC/lib15/tmcent.f:1
c The particles move to the device at the first of masset, findf, step
c   or tmcent after greenm (from ptcls and its zone lists, as restrt or
c   loadup left them) and stay there; after a step that ends on an
c   analysis interval (mod( istep, iphys ) = 0), after tmcent and after
c   masset they are copied back into ptcls with the zone lists rebuilt,
c   and the zone masses into grdmss, grdmoz, jrad and krad, so that
c   densty, dumpp and the other host analysis see the current state.
c   After a solve with the potential the field is copied into grdfld as
c   well.
c Failures call crash( name, message ) with the library's message, as
c   the reference's own routines do (SRC/lib15/crash.f:23-37).
c Set-ups the library does not cover stop in crash: more than one grid,
c   hybrid grids, guard zones, heavy particles, perturbers, unequal
c   masses, recentring, forces on off-grid particles, and anything other
c   than the retarded-velocity leap-frog.
c
      module galhip_state
      use iso_c_binding
      implicit none
c the library context and whether the particles are on the device
      type( c_ptr ), save :: ctx = c_null_ptr
      logical, save :: ondev = .false.
      logical, save :: three = .false.
c ptcls offset of every particle in load order, and the host buffers
      integer( c_int64_t ), save :: nptot = 0
      integer, allocatable, save :: ploct( : )
      integer( c_int32_t ), allocatable, target, save :: pzon( : )
      real( c_float ), allocatable, target, save :: pcrd( :, : )
      end module galhip_state
c
      subroutine galhip_crash( name )
c stops through the reference's crash with the library's last message
      use galhip_f
      implicit none
      character*( * ) name
      character( kind = c_char ), pointer :: msg( : )
      character*200 line
      integer i
      call c_f_pointer( galhip_last_error(), msg, [ 200 ] )
      line = ' '
      do i = 1, 200
        if( msg( i ) .eq. c_null_char )exit
        line( i:i ) = msg( i )
      end do
      call crash( name, trim( line ) )
      end
c
      subroutine galhip_open
c creates the context from the COMMON blocks (p2dset/p3dset's grid,
c   filset's filter, setrun's zones and scales) and the fixed radial
c   force table from the model's force functions
      use galhip_f
      use galhip_state
      implicit none
      include 'inc/params.f'
      include 'inc/admin.f'
      include 'inc/anlsis.f'
      include 'inc/grids.f'
c external
      real axifrc, frcorr, halfrc
c local variables
      type( galhip_p2d_params ) p
      type( galhip_p3d_params ) q
      integer i, k, m, nsk, sk( GALHIP_MAXSKIP )
      integer, parameter :: ntab = 16384
      real*8, allocatable :: fv( : ), fd( : )
      real*8 dr, rmax
      real r
c
      if( c_associated( ctx ) )return
      if( ngrid .ne. 1 )call crash( 'GALHIP', 'one grid only' )
      call switch( 1 )
      if( .not. ( p2d .or. p3d ) )call crash( 'GALHIP',
     +                            'p2d and p3d grids only' )
      if( hybrid .or. heavies .or. pertbn .or. uqmass .or.
     +    ( nguard .gt. 0 ) .or. centrd .or. offfor .or.
     +    ( .not. lfrv ) )call crash( 'GALHIP',
     +                       'set-up not covered by the GPU library' )
      three = p3d
c harmonics the filter removes beyond the sector rule (lg( m + 1 )
c   true = harmonic m skipped, filset.f)
      nsk = 0
      do m = 0, ng - 1
        if( lg( m + 1 ) .and. ( mod( m, nsect ) .eq. 0 ) )then
          if( nsk .eq. GALHIP_MAXSKIP )call crash( 'GALHIP',
     +                               'too many skipped harmonics' )
          nsk = nsk + 1
          sk( nsk ) = m
        end if
      end do
      if( p2d )then
        p%nr = nr( 1 )
        p%na = na
        p%hash = ng - 1
        p%nsect = nsect
        p%nskip = nsk
        p%skip = 0
        p%skip( 1:nsk ) = sk( 1:nsk )
        p%softl = softl
        p%tsoft = tsoft
        p%nzones = nzones
        p%nstep = 1
        p%zbr = 0
        p%nstep( 1:nzones ) = nstep( 1:nzones )
        p%zbr( 1:nzones ) = zbr( 1:nzones )
        p%offanal = merge( 1, 0, offanal )
        p%pmass = pmass
        p%precision = GALHIP_FP32
        i = galhip_p2dset( p, ctx )
        if( i .ne. 0 )call galhip_crash( 'P2DSET' )
        rmax = 1.5d0 * rgrid( 1 )
      else
        q%nr = nr( 1 )
        q%na = na
        q%ngz = ngz
        q%hash = ng - 1
        q%nsect = nsect
        q%nskip = nsk
        q%skip = 0
        q%skip( 1:nsk ) = sk( 1:nsk )
        q%dzg = dzg
        q%softl = softl
        q%tsoft = tsoft
        q%nzones = nzones
        q%nstep = 1
        q%zbr = 0
        q%nstep( 1:nzones ) = nstep( 1:nzones )
        q%zbr( 1:nzones ) = zbr( 1:nzones )
        q%offanal = merge( 1, 0, offanal )
        q%pmass = pmass
        q%precision = GALHIP_FP32
        q%jmass = jmass
        i = galhip_p3dset( q, ctx )
        if( i .ne. 0 )call galhip_crash( 'P3DSET' )
        rmax = 1.5d0 * sqrt( dble( rgrid( 1 ) )**2 +
     +                       dble( zm( 1 ) )**2 )
      end if
c the radial force getacc adds (getacc.f:115-135), tabulated for the
c   device's cubic Hermite interpolation with central-difference slopes
      if( fixrad .or. rigidh .or. suppl )then
        allocate ( fv( 0:ntab - 1 ), fd( 0:ntab - 1 ) )
        dr = rmax / dble( ntab - 1 )
        do k = 1, ntab - 1
          r = real( dr * dble( k ) )
          if( fixrad )then
            fv( k ) = axifrc( r )
          else
            fv( k ) = 0
            if( rigidh )fv( k ) = halfrc( r )
            if( suppl )fv( k ) = fv( k ) + frcorr( r )
          end if
        end do
        fv( 0 ) = 2.d0 * fv( 1 ) - fv( 2 )
        fd( 0 ) = ( fv( 1 ) - fv( 0 ) ) / dr
        fd( ntab - 1 ) = ( fv( ntab - 1 ) - fv( ntab - 2 ) ) / dr
        do k = 1, ntab - 2
          fd( k ) = ( fv( k + 1 ) - fv( k - 1 ) ) / ( 2.d0 * dr )
        end do
        i = galhip_set_fixforce( ctx, ntab, dr, fv, fd )
        if( i .ne. 0 )call galhip_crash( 'GETACC' )
        deallocate ( fv, fd )
      end if
      call switch( 0 )
      return
      end
c
      subroutine galhip_load
c copies every particle of every list from ptcls to the device, with the
c   step counter and zone time stamps, and assigns the mass
      use aarrays
      use galhip_f
      use galhip_state
      implicit none
      include 'inc/params.f'
      include 'inc/admin.f'
c local variables
      integer i, ia, il, j, k, nc
      integer( c_int32_t ), target :: lst( mzones, 2 )
      real a
      equivalence ( ia, a )
c
      if( ondev )return
      call galhip_open
      nc = 4
      if( three )nc = 6
c count the particles on the lists
      nptot = 0
      do il = 1, nlists
        j = islist( 1, il, 1 )
        do while ( j .ge. 0 )
          nptot = nptot + 1
          a = ptcls( j + nwpp )
          j = ia
        end do
      end do
      if( allocated( ploct ) )deallocate ( ploct, pzon, pcrd )
      allocate ( ploct( nptot ), pzon( nptot ), pcrd( nptot, 6 ) )
c list ilist holds zone ilist; the last one the particles off the grid
c   (interpret.f, with one grid and no guard zones)
      k = 0
      do il = 1, nlists
        j = islist( 1, il, 1 )
        do while ( j .ge. 0 )
          k = k + 1
          ploct( k ) = j
          pzon( k ) = min( il, nzones + 1 )
          do i = 1, nc
            pcrd( k, i ) = ptcls( j + i )
          end do
          a = ptcls( j + nwpp )
          j = ia
        end do
      end do
      if( three )then
        i = galhip_restrt3( ctx, nptot, c_loc( pcrd( 1, 1 ) ),
     +      c_loc( pcrd( 1, 2 ) ), c_loc( pcrd( 1, 3 ) ),
     +      c_loc( pcrd( 1, 4 ) ), c_loc( pcrd( 1, 5 ) ),
     +      c_loc( pcrd( 1, 6 ) ), pzon )
      else
        i = galhip_restrt( ctx, nptot, c_loc( pcrd( 1, 1 ) ),
     +      c_loc( pcrd( 1, 2 ) ), c_loc( pcrd( 1, 3 ) ),
     +      c_loc( pcrd( 1, 4 ) ), pzon )
      end if
      if( i .ne. 0 )call galhip_crash( 'RESTRT' )
      lst = 0
      do i = 2, nzones
        lst( i, 1 ) = lstep( 1, i )
        lst( i, 2 ) = lstep( 2, i )
      end do
      if( galhip_set_step( ctx, istep, c_loc( lst ) ) .ne. 0 )
     +                                    call galhip_crash( 'RESTRT' )
      if( galhip_masset( ctx ) .ne. 0 )call galhip_crash( 'MASSET' )
      ondev = .true.
      return
      end
c
      subroutine galhip_store
c copies the particles back into their ptcls records and rebuilds the
c   zone lists (each particle prepended to its list, as scattr does),
c   then the zone masses (fixed point x 2^-qbits) and ring ranges
      use aarrays
      use galhip_f
      use galhip_state
      implicit none
      include 'inc/params.f'
      include 'inc/admin.f'
      include 'inc/grids.f'
c local variables
      integer i, ia, il, j, k, nc, z
      integer( c_int32_t ) jr( mzones ), kr( mzones ), qb
      integer( c_int64_t ) noff8
      integer( c_int64_t ), allocatable, target :: mfx( :, : )
      integer( c_int64_t ), allocatable, target :: moz( :, : )
      real*8 sc
      real a
      equivalence ( ia, a )
c
      nc = 4
      if( three )then
        nc = 6
        i = galhip_unload3( ctx, c_loc( pcrd( 1, 1 ) ),
     +      c_loc( pcrd( 1, 2 ) ), c_loc( pcrd( 1, 3 ) ),
     +      c_loc( pcrd( 1, 4 ) ), c_loc( pcrd( 1, 5 ) ),
     +      c_loc( pcrd( 1, 6 ) ), c_loc( pzon ) )
      else
        i = galhip_unload( ctx, c_loc( pcrd( 1, 1 ) ),
     +      c_loc( pcrd( 1, 2 ) ), c_loc( pcrd( 1, 3 ) ),
     +      c_loc( pcrd( 1, 4 ) ), c_loc( pzon ) )
      end if
      if( i .ne. 0 )call galhip_crash( 'UNLOAD' )
      do il = 1, nlists
        islist( 1, il, 1 ) = -1
      end do
      do k = 1, int( nptot )
        j = ploct( k )
        do i = 1, nc
          ptcls( j + i ) = pcrd( k, i )
        end do
        il = min( int( pzon( k ) ), nlists )
        ia = islist( 1, il, 1 )
        ptcls( j + nwpp ) = a
        islist( 1, il, 1 ) = j
      end do
      if( galhip_get_noff( ctx, noff8 ) .ne. 0 )
     +                                     call galhip_crash( 'UNLOAD' )
      noff = int( noff8 )
      allocate ( mfx( mesh( 1 ), nzones ), moz( mesh( 1 ), nzones ) )
      i = galhip_get_mass( ctx, c_loc( mfx ), c_loc( moz ), jr, kr, qb )
      if( i .ne. 0 )call galhip_crash( 'MASSCL' )
      sc = 2.d0**( -qb )
      do z = 1, nzones
        jrad( z ) = jr( z )
        krad( z ) = kr( z )
        do i = 1, mesh( 1 )
          grdmss( i, z ) = real( dble( mfx( i, z ) ) * sc )
        end do
        if( z .gt. 1 )then
          do i = 1, mesh( 1 )
            grdmoz( i, z ) = real( dble( moz( i, z ) ) * sc )
          end do
        end if
      end do
      deallocate ( mfx, moz )
      return
      end
c
      subroutine galhip_field
c copies the Cartesian field (and potential) of the last solve into
c   grdfld
      use aarrays
      use galhip_f
      use galhip_state
      implicit none
      include 'inc/params.f'
      include 'inc/grids.f'
      integer i, nt
      real( c_float ), allocatable, target :: f( :, : )
      nt = 3
      if( three )nt = 4
      allocate ( f( mesh( 1 ), nt ) )
      if( galhip_get_field( ctx, c_loc( f ) ) .ne. 0 )
     +                                     call galhip_crash( 'FINDF' )
      do i = 1, nt
        grdfld( 1:mesh( 1 ), i ) = f( :, i )
      end do
      deallocate ( f )
      return
      end
c
      subroutine galhip_sync_step
c the step counter, mzone and the zone time stamps from the device
      use galhip_f
      use galhip_state
      implicit none
      include 'inc/params.f'
      include 'inc/admin.f'
      integer i
      integer( c_int32_t ) is, mz
      integer( c_int32_t ), target :: lst( mzones, 2 )
      lst = 0
      i = galhip_get_step( ctx, is, mz, c_loc( lst ) )
      if( i .ne. 0 )call galhip_crash( 'STEP' )
      istep = is
      mzone = mz
      do i = 2, nzones
        lstep( 1, i ) = lst( i, 1 )
        lstep( 2, i ) = lst( i, 2 )
      end do
      return
      end
c
c ---- the overrides -------------------------------------------------
c
      subroutine greenm( direct )
c Green's tables on the device in FP64 (greenm.f:1); the direct kernels
c   of the field test are not formed
      use galhip_f
      use galhip_state
      implicit none
      logical direct
      call galhip_open
      if( galhip_greenm( ctx ) .ne. 0 )call galhip_crash( 'GREENM' )
      return
      end
c
      subroutine masset
      use galhip_f
      use galhip_state
      implicit none
      if( .not. ondev )then
        call galhip_load
      else if( galhip_masset( ctx ) .ne. 0 )then
        call galhip_crash( 'MASSET' )
      end if
      call galhip_store
      return
      end
c
      subroutine findf( lpot )
      use galhip_f
      use galhip_state
      implicit none
      include 'inc/params.f'
      include 'inc/admin.f'
      logical lpot
      call galhip_load
      potl = lpot
      if( galhip_findf( ctx, merge( 1, 0, lpot ) ) .ne. 0 )
     +                                     call galhip_crash( 'FINDF' )
      if( lpot )call galhip_field
      isfld = istep
      return
      end
c
      subroutine p2fndf( jzone, kzone )
c the solve over findf's zone range, polcat included
      use galhip_f
      use galhip_state
      implicit none
      include 'inc/params.f'
      include 'inc/admin.f'
      integer jzone, kzone
      call galhip_load
      if( galhip_p2fndf( ctx, jzone, kzone, merge( 1, 0, potl ) )
     +    .ne. 0 )call galhip_crash( 'P2FNDF' )
      return
      end
c
      subroutine p3fndf( jzone, kzone )
      use galhip_f
      use galhip_state
      implicit none
      include 'inc/params.f'
      include 'inc/admin.f'
      integer jzone, kzone
      call galhip_load
      if( galhip_p2fndf( ctx, jzone, kzone, merge( 1, 0, potl ) )
     +    .ne. 0 )call galhip_crash( 'P3FNDF' )
      return
      end
c
      subroutine polcat
c the device's solve ends with polcat's rotation
      use galhip_state
      implicit none
      if( .not. ondev )call crash( 'POLCAT', 'no device solve yet' )
      return
      end
c
      subroutine step
      use galhip_f
      use galhip_state
      implicit none
      include 'inc/params.f'
      include 'inc/admin.f'
      call galhip_load
      if( galhip_step( ctx ) .ne. 0 )call galhip_crash( 'STEP' )
      call galhip_sync_step
      if( mod( istep, iphys ) .eq. 0 )call galhip_store
      return
      end
c
      subroutine tmcent( start )
      use galhip_f
      use galhip_state
      implicit none
      include 'inc/params.f'
      include 'inc/admin.f'
      logical start
      call galhip_load
      if( galhip_tmcent( ctx, merge( 1, 0, start ) ) .ne. 0 )
     +                                    call galhip_crash( 'TMCENT' )
      tcent = start
      call galhip_store
      return
      end
