! ref_bindings.f90 -- C-callable wrappers around the reference's own
! MPI-free leaf routines (closest_wall.f90, radface.f90, thetaface.f90,
! phiface.f90, findangle3.f90, check.f90, nr_2.f:locate, random.f90:bbfreq/ran,
! ecuyer_cote_mod.f90 via user_set_generator.f90, mathsubs.f90:Bnu/dBnu/expm1,
! nr_2.f:qsimp, zerod.f90).
!
! TEST INFRASTRUCTURE ONLY: linked with the reference sources (compiled
! unmodified from where they lie) into oracle/_ref/libref.so, which the tests
! use to pin the C restatement in oracle/mcrt_oracle.c.  Nothing here is part
! of the product.
module ref_bindings
  use iso_c_binding
  use closest_wall, only : reset_t, find_next_wall, tn1, tp1, tp2, in1, ip1, ip2
  use random, only : ran, bbfreq
  use user_set_generator, only : user_set_all
  implicit none
  real(8), save :: s_k0, s_k1, s_nu0, s_dnu, s_teff   ! integrand state of ref_qsimp_seg
  external qsimp
contains

  subroutine ref_open_log() bind(C, name='ref_open_log')
    ! random.f90:55 writes to unit 12 the first time bbfreq runs
    open(unit=12, file='/dev/null', status='unknown')
  end subroutine ref_open_log

  subroutine ref_walls_state(t, i) bind(C, name='ref_walls_state')
    real(c_double), intent(out) :: t(3)
    integer(c_int), intent(out) :: i(3)
    t(1) = tn1; t(2) = tp1; t(3) = tp2
    i(1) = in1; i(2) = ip1; i(3) = ip2
  end subroutine ref_walls_state

  subroutine ref_reset_t() bind(C, name='ref_reset_t')
    call reset_t()
  end subroutine ref_reset_t

  subroutine ref_radface(u, x, r2a, r2b) bind(C, name='ref_radface')
    real(c_double), intent(in) :: u(3), x(3)
    real(c_double), value :: r2a, r2b
    real(8) :: ux, uy, uz, x1, y1, z1, a, b
    ux = u(1); uy = u(2); uz = u(3); x1 = x(1); y1 = x(2); z1 = x(3); a = r2a; b = r2b
    call radface(ux, uy, uz, x1, y1, z1, a, b)
  end subroutine ref_radface

  subroutine ref_thetaface(u, x, t2a, t2b) bind(C, name='ref_thetaface')
    real(c_double), intent(in) :: u(3), x(3)
    real(c_double), value :: t2a, t2b
    real(8) :: ux, uy, uz, x1, y1, z1, a, b
    ux = u(1); uy = u(2); uz = u(3); x1 = x(1); y1 = x(2); z1 = x(3); a = t2a; b = t2b
    call thetaface(ux, uy, uz, x1, y1, z1, a, b)
  end subroutine ref_thetaface

  subroutine ref_phiface(a1, b1, a2, b2, u, x) bind(C, name='ref_phiface')
    real(c_double), value :: a1, b1, a2, b2
    real(c_double), intent(in) :: u(3), x(3)
    real(8) :: aa1, bb1, aa2, bb2, zero, ux, uy, uz, x1, y1, z1
    aa1 = a1; bb1 = b1; aa2 = a2; bb2 = b2; zero = 0.d0
    ux = u(1); uy = u(2); uz = u(3); x1 = x(1); y1 = x(2); z1 = x(3)
    call phiface(aa1, bb1, zero, zero, aa2, bb2, zero, zero, ux, uy, uz, x1, y1, z1)
  end subroutine ref_phiface

  ! find_next_wall stops the program when no root exists; callers only pass
  ! states with at least one valid candidate.
  subroutine ref_find_next_wall(on_wall, t, i) bind(C, name='ref_find_next_wall')
    integer(c_int), value :: on_wall
    real(c_double), intent(out) :: t
    integer(c_int), intent(out) :: i
    logical(1) :: ow
    real(8) :: tt
    integer :: ii
    ow = (on_wall /= 0)
    call find_next_wall(ow, tt, ii)
    t = tt; i = ii
  end subroutine ref_find_next_wall

  subroutine ref_findangle3(a, out) bind(C, name='ref_findangle3')
    real(c_double), intent(in) :: a(10)
    real(c_double), intent(out) :: out(4)
    real(8) :: c1, s1, p1, cp1, sp1, c2, s2, p2, cp2, sp2, cn, sn, pn, ri1
    c1 = a(1); s1 = a(2); p1 = a(3); cp1 = a(4); sp1 = a(5)
    c2 = a(6); s2 = a(7); p2 = a(8); cp2 = a(9); sp2 = a(10)
    ri1 = 0.d0
    call findangle3(c1, s1, p1, cp1, sp1, c2, s2, p2, cp2, sp2, cn, sn, pn, ri1)
    out(1) = cn; out(2) = sn; out(3) = pn; out(4) = ri1
  end subroutine ref_findangle3

  real(c_double) function ref_check(x) bind(C, name='ref_check')
    real(c_double), value :: x
    real(8) :: check, xx
    character(len=20) :: routine
    external check
    routine = 'ref_bindings'
    xx = x
    ref_check = check(xx, routine)
  end function ref_check

  integer(c_int) function ref_locate(xx, n, x) bind(C, name='ref_locate')
    integer(c_int), value :: n
    real(c_double), intent(in) :: xx(n)
    real(c_double), value :: x
    real(8) :: arr(n), xv
    integer :: j, nn
    arr = xx; xv = x; nn = n
    call locate(arr, nn, xv, j)
    ref_locate = j
  end function ref_locate

  subroutine ref_user_set_all(s1, s2, g) bind(C, name='ref_user_set_all')
    integer(c_int), value :: s1, s2, g
    call user_set_all(int(s1), int(s2), int(g))
  end subroutine ref_user_set_all

  real(c_double) function ref_ran() bind(C, name='ref_ran')
    ref_ran = ran()
  end function ref_ran

  real(c_double) function ref_bbfreq(teff) bind(C, name='ref_bbfreq')
    real(c_double), value :: teff
    real(8) :: t
    t = teff
    ref_bbfreq = bbfreq(t)
  end function ref_bbfreq

  real(c_double) function ref_bnu(nu, t) bind(C, name='ref_bnu')
    real(c_double), value :: nu, t
    real(8) :: Bnu, a, b
    external Bnu
    a = nu; b = t
    ref_bnu = Bnu(a, b)
  end function ref_bnu

  real(c_double) function ref_dbnu(nu, t) bind(C, name='ref_dbnu')
    real(c_double), value :: nu, t
    real(8) :: dBnu, a, b
    external dBnu
    a = nu; b = t
    ref_dbnu = dBnu(a, b)
  end function ref_dbnu

  real(c_double) function ref_expm1(x) bind(C, name='ref_expm1')
    real(c_double), value :: x
    real(8) :: expm1, a
    external expm1
    a = x
    ref_expm1 = expm1(a)
  end function ref_expm1

  ! One interval of setemiss (dust_5.f90:560-575): the reference's qsimp
  ! (nr_2.f:209-231) over f in [0,1] of kappaB (kind 0) or chidB (kind 1).
  ! The two integrands are dust_5.f90:782-794 and :814-826 restated (that
  ! file needs mpif.h); Bnu, dBnu, expm1 and qsimp are the reference's.
  subroutine ref_qsimp_seg(kind, k0, k1, nu0, dnu, teff, s, converg) bind(C, name='ref_qsimp_seg')
    integer(c_int), value :: kind
    real(c_double), value :: k0, k1, nu0, dnu, teff
    real(c_double), intent(out) :: s
    integer(c_int), intent(out) :: converg
    real(8) :: a, b, r
    logical :: cv
    s_k0 = k0; s_k1 = k1; s_nu0 = nu0; s_dnu = dnu; s_teff = teff
    a = 0.d0; b = 1.d0
    if (kind == 0) then
       call qsimp(seg_kappab, a, b, r, cv)
    else
       call qsimp(seg_chidb, a, b, r, cv)
    end if
    s = r
    converg = merge(1, 0, cv)
  end subroutine ref_qsimp_seg

  real(8) function seg_kappab(f)
    real(8) :: f, nu, Bnu
    external Bnu
    nu = s_nu0 + f * s_dnu
    seg_kappab = ((1.d0 - f) * s_k0 + f * s_k1) * Bnu(nu, s_teff)
  end function seg_kappab

  real(8) function seg_chidb(f)
    real(8) :: f, nu, dBnu
    external dBnu
    nu = s_nu0 + f * s_dnu
    seg_chidb = dBnu(nu, s_teff) / ((1.d0 - f) * s_k0 + f * s_k1)
  end function seg_chidb

  ! zerod.f90: the Ulrich streamline cubic.  xmu0 is written only when the
  ! routine finds its root; iflag is 1 in its "2 roots" case.
  subroutine ref_zerod(xmu, z, xmu0, iflag) bind(C, name='ref_zerod')
    real(c_double), value :: xmu, z
    real(c_double), intent(inout) :: xmu0
    integer(c_int), intent(out) :: iflag
    real(8) :: a, b, r
    integer :: fl
    a = xmu; b = z; r = xmu0; fl = 0
    call zerod(a, b, r, fl)
    xmu0 = r
    iflag = fl
  end subroutine ref_zerod

end module ref_bindings
