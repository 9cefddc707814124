! disk_gpu.f90 -- drop-in for the packet loop of disk (newdisktherm_5_fuv.f90:1).
!
! The reference's disk runs `do iphot=1,np` over propagate on every MPI rank
! (:220-491), sums the tallies with MPI_ALLREDUCE (:524-602) and calls tfinal
! (:620).  disk_gpu does the same work on one GPU: one mcrt_run_packets call
! for all np packets (counter-based RNG: packet k of iteration `iter` draws
! from stream (seed, k) independently of any rank split), one tally read-back
! (the post-ALLREDUCE state), and the FP64 Lucy inversion on the device,
! with tfinal's convergence test: converge becomes 'yes' when it passes
! (tfinal_5_test1.f90:334-368; newtts.f90:215-256 loops while it is 'no').
! Array shapes are the reference's module arrays:
!   dtauabs(nrg,ntg,npg) real(8), nabs(nrg,ntg,npg) integer, tdave(nrg,ntg,npg)
!   sed(nfreq,nmu,nph,nap,8,10)  = si,sq,su,sv,si2,sq2,su2,sv2,nums,aveinc
!   psed(nfreq,npeel,nap,8,9)    = ti,tq,tu,tv,ti2,tq2,tu2,tv2,numt
!   image(npeel,nx,nx,nband,8)   = image_b_{i,q,u,v,d,s,e,o}
subroutine disk_gpu(converge, ctx, np, seed, iter, precision, tmode, nrg, ntg, npg, dtauabs, nabs, tdave, &
     nfreq, nmu, nph, nap, npeel, nx, nband, sed, psed, image, counters)
  use iso_c_binding
  use mcrt_mod
  implicit none
  character(len=3), intent(inout) :: converge
  type(c_ptr), intent(in) :: ctx
  integer(8), intent(in) :: np, seed
  integer, intent(in) :: iter, precision, tmode, nrg, ntg, npg
  integer, intent(in) :: nfreq, nmu, nph, nap, npeel, nx, nband
  real(8), intent(inout) :: dtauabs(nrg, ntg, npg), tdave(nrg, ntg, npg)
  integer, intent(inout) :: nabs(nrg, ntg, npg)
  real(4), intent(inout), target :: sed(nfreq, nmu, nph, nap, 8, 10)
  real(4), intent(inout), target :: psed(nfreq, max(npeel, 1), nap, 8, 9)
  real(4), intent(inout), target :: image(max(npeel, 1), nx, nx, max(nband, 1), 8)
  real(8), intent(out), target :: counters(16)
  real(8), allocatable, target :: dt(:,:,:), td(:,:,:)
  integer(c_int32_t), allocatable, target :: na(:,:,:)
  real(c_double) :: sig, sigprev
  integer(c_int64_t) :: ncount
  integer(c_int32_t) :: conv

  allocate(dt(nrg-1, ntg-1, npg-1), td(nrg-1, ntg-1, npg-1), na(nrg-1, ntg-1, npg-1))
  call mcrt_check(ctx, mcrt_zero_tallies(ctx), 'mcrt_zero_tallies')
  ! packets of iteration `iter` use counters iter*np .. iter*np+np-1
  call mcrt_check(ctx, mcrt_run_packets(ctx, int(iter, c_int64_t) * np, np, seed, &
       int(precision, c_int32_t)), 'mcrt_run_packets')
  call mcrt_check(ctx, mcrt_get_tallies(ctx, c_loc(dt), c_loc(na), c_null_ptr, c_null_ptr, &
       c_loc(sed), c_loc(psed), c_loc(image), c_loc(counters)), 'mcrt_get_tallies')
  call mcrt_check(ctx, mcrt_dusttemp(ctx, real(np, c_double), int(tmode, c_int32_t), 1_c_int32_t, &
       c_loc(td)), 'mcrt_dusttemp')
  call mcrt_check(ctx, mcrt_dusttemp_sigma(ctx, sig, sigprev, ncount, conv), 'mcrt_dusttemp_sigma')
  if (converge == 'no' .and. conv /= 0) converge = 'yes'
  dtauabs(1:nrg-1, 1:ntg-1, 1:npg-1) = dt
  nabs(1:nrg-1, 1:ntg-1, 1:npg-1) = na
  tdave(1:nrg-1, 1:ntg-1, 1:npg-1) = td
  deallocate(dt, td, na)
end subroutine disk_gpu
