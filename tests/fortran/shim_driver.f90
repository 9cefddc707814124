! shim_driver.f90 -- test driver for the Fortran host shim.  Reads a model
! dumped by tests/test_gpu_fortran_shim.py in the reference's module layout
! (walls, vcell/densarr/findopac/tdust over (nrg,ntg,npg[,10])), uploads it
! through mcrt_mod, runs one disk_gpu iteration and writes the tallies back.
program shim_driver
  use iso_c_binding
  use mcrt_mod
  implicit none
  character(len=512) :: inpath, outpath
  type(c_ptr) :: ctx
  integer :: u, nrg, ntg, npg, nopac, nlam, nT, nnu, npeel, nx, nband, nwmax
  integer :: nfreq, nmu, nph, nap, precision, iter
  integer(8) :: np, seed
  real(8) :: rstar, tstar, rmin, rmaxi, numin, lnurat
  real(8), allocatable :: rarr(:), r2arr(:), thetarr(:), costarr(:), tan2arr(:), phiarr(:), aarr(:), barr(:)
  real(8), allocatable :: vcell(:,:,:), densarr(:,:,:,:), tdust(:,:,:), dtauabs(:,:,:), tdave(:,:,:)
  integer, allocatable :: findopac(:,:,:,:), nabs(:,:,:), band_nw(:)
  real(8), allocatable :: nudust(:), kappad(:), adust(:), gdust(:), pldust(:), kappav(:)
  real(8), allocatable :: lTint(:), Tint(:), nuint(:), kapint(:), kappap(:)
  real(8), allocatable :: thete(:), phie(:), band_freq(:), band_phi(:), aperture2(:), u_mu(:)
  real(4), allocatable :: sed(:,:,:,:,:,:), psed(:,:,:,:,:), image(:,:,:,:,:)
  real(8) :: counters(16)
  integer(c_int32_t), allocatable :: bnw(:)
  character(len=3) :: converge

  call get_command_argument(1, inpath)
  call get_command_argument(2, outpath)
  open(newunit=u, file=trim(inpath), access='stream', form='unformatted', status='old')
  read(u) nrg, ntg, npg, nopac, nlam, nT, nnu, npeel, nx, nband, nwmax, nfreq, nmu, nph, nap, precision, iter
  read(u) np, seed
  read(u) rstar, tstar, rmin, rmaxi, numin, lnurat
  allocate(rarr(nrg), r2arr(nrg), thetarr(ntg), costarr(ntg), tan2arr(ntg), phiarr(npg), aarr(npg), barr(npg))
  read(u) rarr, r2arr, thetarr, costarr, tan2arr, phiarr, aarr, barr
  allocate(vcell(nrg, ntg, npg), densarr(nrg, ntg, npg, 10), findopac(nrg, ntg, npg, 10), tdust(nrg, ntg, npg))
  read(u) vcell, densarr, findopac, tdust
  allocate(nudust(nopac*nlam), kappad(nopac*nlam), adust(nopac*nlam), gdust(nopac*nlam), pldust(nopac*nlam))
  allocate(kappav(nopac), lTint(nT), Tint(nT), nuint(nnu), kapint(nopac*nnu*nT), kappap(nopac*nT))
  read(u) nudust, kappad, adust, gdust, pldust, kappav, lTint, Tint, nuint, kapint, kappap
  allocate(thete(npeel), phie(npeel), band_nw(nband), band_freq(nband*nwmax), band_phi(nband*nwmax))
  allocate(aperture2(nap), u_mu(nmu))
  read(u) thete, phie, band_nw, band_freq, band_phi, aperture2, u_mu
  close(u)

  if (mcrt_create(0_c_int, ctx) /= 0) then
    print *, 'mcrt_create failed'
    stop 2
  end if
  call mcrt_put_grid(ctx, nrg, ntg, npg, rarr, r2arr, thetarr, costarr, tan2arr, phiarr, aarr, barr, vcell)
  call mcrt_put_medium(ctx, nrg, ntg, npg, densarr, findopac)
  call mcrt_check(ctx, mcrt_set_dust(ctx, int(nopac, c_int32_t), int(nlam, c_int32_t), nudust, kappad, adust, &
       gdust, pldust, kappav, int(nT, c_int32_t), int(nnu, c_int32_t), lTint, Tint, nuint, kapint, kappap), &
       'mcrt_set_dust')
  call mcrt_check(ctx, mcrt_set_source(ctx, rstar, tstar, rmin), 'mcrt_set_source')
  allocate(bnw(nband))
  bnw = int(band_nw, c_int32_t)
  call mcrt_check(ctx, mcrt_set_observer(ctx, 1_c_int32_t, int(npeel, c_int32_t), thete, phie, &
       int(nx, c_int32_t), rmaxi, int(nband, c_int32_t), int(nwmax, c_int32_t), bnw, band_freq, band_phi, &
       int(nfreq, c_int32_t), int(nmu, c_int32_t), int(nph, c_int32_t), int(nap, c_int32_t), aperture2, u_mu, &
       numin, lnurat), 'mcrt_set_observer')
  call mcrt_put_tdust(ctx, nrg, ntg, npg, tdust)

  allocate(dtauabs(nrg, ntg, npg), tdave(nrg, ntg, npg), nabs(nrg, ntg, npg))
  allocate(sed(nfreq, nmu, nph, nap, 8, 10), psed(nfreq, npeel, nap, 8, 9), image(npeel, nx, nx, nband, 8))
  dtauabs = 0.d0; tdave = 0.d0; nabs = 0
  converge = 'no'
  call disk_gpu(converge, ctx, np, seed, iter, precision, MCRT_T_REFERENCE, nrg, ntg, npg, dtauabs, nabs, tdave, &
       nfreq, nmu, nph, nap, npeel, nx, nband, sed, psed, image, counters)
  if (mcrt_destroy(ctx) /= 0) stop 3

  open(newunit=u, file=trim(outpath), access='stream', form='unformatted', status='replace')
  write(u) dtauabs, nabs, tdave, sed, psed, image, counters
  close(u)
  print '(a,f12.0,a,f12.0)', 'shim_driver: escaped ', counters(10), ' errors ', counters(16)
end program shim_driver
