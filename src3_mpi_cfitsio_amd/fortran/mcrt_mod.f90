! mcrt_mod.f90 -- Fortran host shim over libmcrt.so (include/mcrt.h).
!
! The reference keeps its packet-path state in module arrays sized by walls,
! (nrg,ntg,npg) with 1-based indices (grid_mod_5.f90:29-63, dust_mod_5.f90,
! tts_mod_5.f90).  This module binds the C ABI with iso_c_binding and adds
! helpers that take those arrays exactly as the reference holds them:
!   mcrt_put_grid    <- rarr, r2arr, thetarr, costarr, tan2arr, phiarr, aarr,
!                       barr, vcell(nrg,ntg,npg)
!   mcrt_put_medium  <- densarr(nrg,ntg,npg,ndg+2), findopac(nrg,ntg,npg,ndg+2)
!   mcrt_take_cells  -> dtauabs / nabs style (nrg,ntg,npg) arrays
! so a replacement of disk (newdisktherm_5_fuv.f90:1) is a handful of calls;
! see disk_gpu.f90.
module mcrt_mod
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: MCRT_FP64 = 0, MCRT_MIXED = 1
  integer(c_int), parameter :: MCRT_T_REFERENCE = 0, MCRT_T_PERCELL = 1
  integer, parameter :: MCRT_NSPECIES = 10

  interface
    integer(c_int) function mcrt_create(device, ctx) bind(C, name='mcrt_create')
      import :: c_int, c_ptr
      integer(c_int), value :: device
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function mcrt_destroy(ctx) bind(C, name='mcrt_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function mcrt_last_error(ctx) bind(C, name='mcrt_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mcrt_set_grid(ctx, nrg, ntg, npg, rarr, r2arr, thetarr, costarr, tan2arr, &
         phiarr, aarr, barr, vcell) bind(C, name='mcrt_set_grid')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nrg, ntg, npg
      real(c_double), intent(in) :: rarr(*), r2arr(*), thetarr(*), costarr(*), tan2arr(*)
      real(c_double), intent(in) :: phiarr(*), aarr(*), barr(*), vcell(*)
    end function
    integer(c_int) function mcrt_set_medium(ctx, rho, species, species_opac) bind(C, name='mcrt_set_medium')
      import :: c_int, c_ptr, c_int32_t, c_double, c_int8_t
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: rho(*)
      integer(c_int8_t), intent(in) :: species(*)
      integer(c_int32_t), intent(in) :: species_opac(0:10)
    end function
    integer(c_int) function mcrt_set_dust(ctx, nopac, nlam, nudust, kappad, adust, gdust, pldust, kappav, &
         nT, nnu, lTint, Tint, nuint, kapint, kappap) bind(C, name='mcrt_set_dust')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nopac, nlam, nT, nnu
      real(c_double), intent(in) :: nudust(*), kappad(*), adust(*), gdust(*), pldust(*), kappav(*)
      real(c_double), intent(in) :: lTint(*), Tint(*), nuint(*), kapint(*), kappap(*)
    end function
    ! setemiss (dust_5.f90:507-699) on the GPU; arrays in dust_mod's layout
    integer(c_int) function mcrt_setemiss(ctx, nopac, nlam, nlambda, nudust, kappad, adust, nT, Tint, nnu, &
         nuint, smooth, kapint, kappap, kappar, nfail, nevals, ms) bind(C, name='mcrt_setemiss')
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double, c_float
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nopac, nlam, nT, nnu
      integer(c_int32_t), intent(in) :: nlambda(*), smooth(*)
      real(c_double), intent(in) :: nudust(*), kappad(*), adust(*), Tint(*), nuint(*)
      real(c_double), intent(out) :: kapint(*), kappap(*), kappar(*)
      integer(c_int64_t), intent(out) :: nfail, nevals
      real(c_float), intent(out) :: ms
    end function
    ! external illumination: isrf_mod's tables and isrf_frac
    integer(c_int) function mcrt_set_isrf(ctx, nisrf, cdf, logcdf, logfreq, isrf_frac) bind(C, name='mcrt_set_isrf')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nisrf
      real(c_double), intent(in) :: cdf(*), logcdf(*), logfreq(*)
      real(c_double), value :: isrf_frac
    end function
    ! stellar hot spots: spot_mod's parameters and the hot-spot photon fraction
    integer(c_int) function mcrt_set_spots(ctx, nspot, fspot, spotlat, spotlon, tshock, accfrac2) &
         bind(C, name='mcrt_set_spots')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nspot
      real(c_double), value :: fspot, spotlat, spotlon, tshock, accfrac2
    end function
    ! Ldint and hdisk of grid_mod (nrg-1 each), accfrac of tts_mod
    integer(c_int) function mcrt_set_accretion(ctx, ldint, hdisk, accfrac) bind(C, name='mcrt_set_accretion')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: ldint(*), hdisk(*)
      real(c_double), value :: accfrac
    end function
    ! diffusion approximation of thick regions (setdiffus_5.f90, emit_common_2.f90)
    integer(c_int) function mcrt_set_diffusion(ctx, on, kappar, hdisk, rmaxd, nr0) &
         bind(C, name='mcrt_set_diffusion')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on, nr0
      real(c_double), intent(in) :: kappar(*), hdisk(*)
      real(c_double), value :: rmaxd
    end function
    integer(c_int) function mcrt_setdiffus(ctx, iter, ncells) bind(C, name='mcrt_setdiffus')
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: iter
      integer(c_int64_t), intent(out) :: ncells
    end function
    integer(c_int) function mcrt_get_diffusion(ctx, itmin, count) bind(C, name='mcrt_get_diffusion')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: itmin(*), count(*)
    end function
    integer(c_int) function mcrt_set_source(ctx, rstar, tstar, rmin) bind(C, name='mcrt_set_source')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: rstar, tstar, rmin
    end function
    integer(c_int) function mcrt_set_observer(ctx, ipeel, npeel, thete, phie, nx, rmaxi, nband, nwmax, &
         band_nw, band_freq, band_phi, nfreq, nmu, nph, nap, aperture2, u_mu, numin, lnurat) &
         bind(C, name='mcrt_set_observer')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ipeel, npeel, nx, nband, nwmax, nfreq, nmu, nph, nap
      real(c_double), intent(in) :: thete(*), phie(*), band_freq(*), band_phi(*), aperture2(*), u_mu(*)
      integer(c_int32_t), intent(in) :: band_nw(*)
      real(c_double), value :: rmaxi, numin, lnurat
    end function
    integer(c_int) function mcrt_set_tdust(ctx, tdust) bind(C, name='mcrt_set_tdust')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: tdust(*)
    end function
    integer(c_int) function mcrt_zero_tallies(ctx) bind(C, name='mcrt_zero_tallies')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mcrt_run_packets(ctx, first_packet, npackets, seed, precision) &
         bind(C, name='mcrt_run_packets')
      import :: c_int, c_ptr, c_int64_t, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: first_packet, npackets, seed
      integer(c_int32_t), value :: precision
    end function
    integer(c_int) function mcrt_get_tallies(ctx, dtauabs, nabs, dbig, killed, sed, psed, image, counters) &
         bind(C, name='mcrt_get_tallies')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, dtauabs, nabs, dbig, killed, sed, psed, image, counters
    end function
    integer(c_int) function mcrt_dusttemp(ctx, nphot, mode, update_tdust, tdave) bind(C, name='mcrt_dusttemp')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: nphot
      integer(c_int32_t), value :: mode, update_tdust
      type(c_ptr), value :: tdave
    end function
    ! tfinal's convergence test of the last mcrt_dusttemp (tfinal_5_test1.f90:334-368)
    integer(c_int) function mcrt_dusttemp_sigma(ctx, sigma, sigma_prev, ncells, converge) &
         bind(C, name='mcrt_dusttemp_sigma')
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: sigma, sigma_prev
      integer(c_int64_t), intent(out) :: ncells
      integer(c_int32_t), intent(out) :: converge
    end function
    ! loss sum(W * image_I) and d loss / d rho per cell (path replay)
    integer(c_int) function mcrt_adjoint_image(ctx, first_packet, npackets, seed, precision, w, loss, grad) &
         bind(C, name='mcrt_adjoint_image')
      import :: c_int, c_ptr, c_int64_t, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: first_packet, npackets, seed
      integer(c_int32_t), value :: precision
      type(c_ptr), value :: w, loss, grad
    end function
  end interface

contains

  ! the library's message for the last failure, as a Fortran string
  function mcrt_error_text(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=256) :: msg
    character(kind=c_char), pointer :: s(:)
    integer :: i
    msg = ' '
    call c_f_pointer(mcrt_last_error(ctx), s, [256])
    do i = 1, 256
      if (s(i) == c_null_char) exit
      msg(i:i) = s(i)
    end do
  end function mcrt_error_text

  ! stop with the library's message, as the reference stops on fatal errors
  subroutine mcrt_check(ctx, rc, where)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= 0) then
      print *, 'mcrt error in ', where, ': ', trim(mcrt_error_text(ctx))
      stop 1
    end if
  end subroutine mcrt_check

  ! wall arrays and the (nrg,ntg,npg) vcell of gridset, cells copied out
  subroutine mcrt_put_grid(ctx, nrg, ntg, npg, rarr, r2arr, thetarr, costarr, tan2arr, phiarr, aarr, barr, vcell)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: nrg, ntg, npg
    real(8), intent(in) :: rarr(nrg), r2arr(nrg), thetarr(ntg), costarr(ntg), tan2arr(ntg)
    real(8), intent(in) :: phiarr(npg), aarr(npg), barr(npg), vcell(nrg, ntg, npg)
    real(8), allocatable :: vc(:,:,:)
    allocate(vc(nrg-1, ntg-1, npg-1))
    vc = vcell(1:nrg-1, 1:ntg-1, 1:npg-1)
    call mcrt_check(ctx, mcrt_set_grid(ctx, int(nrg, c_int32_t), int(ntg, c_int32_t), int(npg, c_int32_t), &
         rarr, r2arr, thetarr, costarr, tan2arr, phiarr, aarr, barr, vc), 'mcrt_set_grid')
    deallocate(vc)
  end subroutine mcrt_put_grid

  ! densarr(nrg,ntg,npg,ndg+2) + findopac -> one thermal species per cell
  ! (select_dust_2.f90:22-34 semantics); findopac values are 1-based opacity ids
  subroutine mcrt_put_medium(ctx, nrg, ntg, npg, densarr, findopac)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: nrg, ntg, npg
    real(8), intent(in) :: densarr(nrg, ntg, npg, MCRT_NSPECIES)
    integer, intent(in) :: findopac(nrg, ntg, npg, MCRT_NSPECIES)
    real(8), allocatable :: rho(:,:,:)
    integer(c_int8_t), allocatable :: sp(:,:,:)
    integer(c_int32_t) :: sop(0:10)
    integer :: ir, it, ip, id, nn
    allocate(rho(nrg-1, ntg-1, npg-1), sp(nrg-1, ntg-1, npg-1))
    rho = 0.d0
    sp = 0_c_int8_t
    sop = 0
    do ip = 1, npg-1
      do it = 1, ntg-1
        do ir = 1, nrg-1
          nn = 0
          do id = 1, MCRT_NSPECIES
            if (densarr(ir, it, ip, id) > 0.d0) then
              nn = nn + 1
              rho(ir, it, ip) = densarr(ir, it, ip, id)
              sp(ir, it, ip) = int(id, c_int8_t)
              sop(id) = int(findopac(ir, it, ip, id) - 1, c_int32_t)
            end if
          end do
          if (nn > 1) then
            print *, 'mcrt_put_medium: more than one non-zero density', ir, it, ip
            stop 1
          end if
        end do
      end do
    end do
    call mcrt_check(ctx, mcrt_set_medium(ctx, rho, sp, sop), 'mcrt_set_medium')
    deallocate(rho, sp)
  end subroutine mcrt_put_medium

  ! reference-layout (nrg,ntg,npg) temperature -> device re-emission temperature
  subroutine mcrt_put_tdust(ctx, nrg, ntg, npg, tdust)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: nrg, ntg, npg
    real(8), intent(in) :: tdust(nrg, ntg, npg)
    real(8), allocatable :: t(:,:,:)
    allocate(t(nrg-1, ntg-1, npg-1))
    t = tdust(1:nrg-1, 1:ntg-1, 1:npg-1)
    call mcrt_check(ctx, mcrt_set_tdust(ctx, t), 'mcrt_set_tdust')
    deallocate(t)
  end subroutine mcrt_put_tdust

  ! cell array from the device into a reference-layout (nrg,ntg,npg) array
  subroutine mcrt_cells_to_walls(nrg, ntg, npg, cells, full)
    integer, intent(in) :: nrg, ntg, npg
    real(8), intent(in) :: cells(nrg-1, ntg-1, npg-1)
    real(8), intent(inout) :: full(nrg, ntg, npg)
    full(1:nrg-1, 1:ntg-1, 1:npg-1) = cells
  end subroutine mcrt_cells_to_walls

end module mcrt_mod
