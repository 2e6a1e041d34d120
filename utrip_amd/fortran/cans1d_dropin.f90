!=======================================================================
! libcans1d_gpu -- the cans1d hot-path routines with the reference's names
! and argument lists (cans1d/hdmlw/mlw_m.f:2, cans1d/hdroe/roe_m.f:2,
! cans1d/common/cfl_m.f:2, cans1d/bc/bd*x.f, cans1d/hdmlw/mlw_m3.f:2,
! cans1d/hdroe/roe_m3.f:2, cans1d/common/cfl_m3.f:2, and for the UTRIP
! spicule model
! cans1d/hdmlw/mlw_m_bg.f:2, cans1d/common/chkdav.f:2,
! cans1d/md_spicule/hdmlw/trqalfven.f:2), computed on the GPU by
! libcans_gpu.so.  A cans1d main program links this library instead of
! libcans1d.a and runs unchanged.
!=======================================================================
subroutine mlw_m(ro, pr, vx, vy, by, bx, bxm, dt, qav, gm, dx, dxm, ix)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix
  double precision, target :: ro(ix), pr(ix), vx(ix), vy(ix), by(ix)
  double precision, target, intent(in) :: bx(ix), bxm(ix), dx(ix), dxm(ix)
  double precision, intent(in) :: dt, qav, gm
  call check(c_mlw1d(dmap(ro, ix), dmap(pr, ix), dmap(vx, ix), dmap(vy, ix), dmap(by, ix), dmap(bx, ix), &
       dmap(bxm, ix), dt, qav, gm, dmap(dx, ix), dmap(dxm, ix), int(ix, c_int), c_null_ptr), 'mlw_m')
  call dunmap(ro, 1); call dunmap(pr, 1); call dunmap(vx, 1); call dunmap(vy, 1); call dunmap(by, 1)
  call dunmap(bx, 0); call dunmap(bxm, 0); call dunmap(dx, 0); call dunmap(dxm, 0)
end subroutine mlw_m

subroutine mlw_m_bg(ro, pr, vx, vy, by, bx, bxm, dt, qav, gm, gx, gxm, sc, dsc, scm, dscm, rr, rrm, drr, drrm, &
     dx, dxm, ix)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix
  double precision, target :: ro(ix), pr(ix), vx(ix), vy(ix), by(ix)
  double precision, target, intent(in) :: bx(ix), bxm(ix), gx(ix), gxm(ix), sc(ix), dsc(ix), scm(ix), dscm(ix)
  double precision, target, intent(in) :: rr(ix), rrm(ix), drr(ix), drrm(ix), dx(ix), dxm(ix)
  double precision, intent(in) :: dt, qav, gm
  call check(c_mlw1d_bg(dmap(ro, ix), dmap(pr, ix), dmap(vx, ix), dmap(vy, ix), dmap(by, ix), dmap(bx, ix), &
       dmap(bxm, ix), dt, qav, gm, dmap(gx, ix), dmap(gxm, ix), dmap(sc, ix), dmap(dsc, ix), dmap(scm, ix), &
       dmap(dscm, ix), dmap(rr, ix), dmap(rrm, ix), dmap(drr, ix), dmap(drrm, ix), dmap(dx, ix), dmap(dxm, ix), &
       int(ix, c_int), c_null_ptr), 'mlw_m_bg')
  call dunmap(ro, 1); call dunmap(pr, 1); call dunmap(vx, 1); call dunmap(vy, 1); call dunmap(by, 1)
  call dunmap(bx, 0); call dunmap(bxm, 0); call dunmap(gx, 0); call dunmap(gxm, 0); call dunmap(sc, 0)
  call dunmap(dsc, 0); call dunmap(scm, 0); call dunmap(dscm, 0); call dunmap(rr, 0); call dunmap(rrm, 0)
  call dunmap(drr, 0); call dunmap(drrm, 0); call dunmap(dx, 0); call dunmap(dxm, 0)
end subroutine mlw_m_bg

subroutine chkdav(n_events, da, vx, floor, ix)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix
  integer, intent(out) :: n_events
  double precision, target :: da(ix), vx(ix)
  double precision, intent(in) :: floor
  integer(c_int) :: n
  call check(c_chkdav1d(n, dmap(da, ix), dmap(vx, ix), floor, int(ix, c_int), c_null_ptr), 'chkdav')
  call dunmap(da, 1); call dunmap(vx, 1)
  n_events = n
end subroutine chkdav

!-----------------------------------------------------------------------
! trqalfven with the Park-Miller stream of cans1d/common/rangen.f in place of
! the compiler's RANDOM_NUMBER (one draw per cell per call, cell order).
! The seed is this library's (cans_gpu_rangen_seed sets it; default 1).
! The torque profile tanh((x-0.75)/0.075)-1 depends on x only: it is
! evaluated on the host (this compiler's libm tanh) when x changes.
!-----------------------------------------------------------------------
module cans_spicule_state
  implicit none
  integer, save :: mseed = 1
  double precision, allocatable, target, save :: trq(:), xsaved(:)
end module cans_spicule_state

subroutine cans_gpu_rangen_seed(iseed)
  use cans_spicule_state
  implicit none
  integer, intent(in) :: iseed
  mseed = iseed
end subroutine cans_gpu_rangen_seed

subroutine trqalfven(vy, dt, time, x, ix, rr)
  use cans_gpu_iface
  use cans_spicule_state
  implicit none
  integer, intent(in) :: ix
  double precision, target :: vy(ix)
  double precision, target, intent(in) :: x(ix), rr(ix)
  double precision, intent(in) :: dt, time
  integer :: i
  integer(c_int) :: s
  logical :: fresh
  fresh = .not. allocated(trq)
  if (.not. fresh) fresh = size(trq) /= ix
  if (.not. fresh) fresh = any(xsaved /= x)
  if (fresh) then
     if (allocated(trq)) then
        call check(c_forget(c_loc(trq(1))), 'trqalfven')
        deallocate(trq, xsaved)
     end if
     allocate(trq(ix), xsaved(ix))
     do i = 1, ix
        trq(i) = tanh((x(i) - 0.75) / 0.075) - 1
     end do
     trq(ix) = trq(ix - 2)
     xsaved = x
  end if
  s = mseed
  call check(c_trq1d(dmap(vy, ix), dt, dmap(trq, ix), dmap(rr, ix), int(ix, c_int), s, c_null_ptr), 'trqalfven')
  mseed = s
  call dunmap(vy, 1); call dunmap(trq, 0); call dunmap(rr, 0)
end subroutine trqalfven

subroutine mlw_m3(ro, pr, vx, vy, vz, by, bz, bx, bxm, dt, qav, gm, dx, dxm, ix)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix
  double precision, target :: ro(ix), pr(ix), vx(ix), vy(ix), vz(ix), by(ix), bz(ix)
  double precision, target, intent(in) :: bx(ix), bxm(ix), dx(ix), dxm(ix)
  double precision, intent(in) :: dt, qav, gm
  call check(c_mlw1d_m3(dmap(ro, ix), dmap(pr, ix), dmap(vx, ix), dmap(vy, ix), dmap(vz, ix), dmap(by, ix), &
       dmap(bz, ix), dmap(bx, ix), dmap(bxm, ix), dt, qav, gm, dmap(dx, ix), dmap(dxm, ix), int(ix, c_int), &
       c_null_ptr), 'mlw_m3')
  call dunmap(ro, 1); call dunmap(pr, 1); call dunmap(vx, 1); call dunmap(vy, 1); call dunmap(vz, 1)
  call dunmap(by, 1); call dunmap(bz, 1)
  call dunmap(bx, 0); call dunmap(bxm, 0); call dunmap(dx, 0); call dunmap(dxm, 0)
end subroutine mlw_m3

subroutine roe_m3(ro, pr, vx, vy, vz, by, bz, bx, bxm, dt, gm, dx, ix)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix
  double precision, target :: ro(ix), pr(ix), vx(ix), vy(ix), vz(ix), by(ix), bz(ix)
  double precision, target, intent(in) :: bx(ix), bxm(ix), dx(ix)
  double precision, intent(in) :: dt, gm
  call check(c_roe1d_m3(dmap(ro, ix), dmap(pr, ix), dmap(vx, ix), dmap(vy, ix), dmap(vz, ix), dmap(by, ix), &
       dmap(bz, ix), dmap(bx, ix), dmap(bxm, ix), dt, gm, dmap(dx, ix), int(ix, c_int), c_null_ptr), 'roe_m3')
  call dunmap(ro, 1); call dunmap(pr, 1); call dunmap(vx, 1); call dunmap(vy, 1); call dunmap(vz, 1)
  call dunmap(by, 1); call dunmap(bz, 1)
  call dunmap(bx, 0); call dunmap(bxm, 0); call dunmap(dx, 0)
end subroutine roe_m3

subroutine cfl_m3(dt, safety, dtmin, merr, gm, bx, ro, pr, vx, vy, by, vz, bz, dx, ix)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix
  integer, intent(out) :: merr
  double precision, intent(out) :: dt
  double precision, intent(in) :: safety, dtmin, gm
  double precision, target, intent(in) :: bx(ix), ro(ix), pr(ix), vx(ix), vy(ix), by(ix), vz(ix), bz(ix), dx(ix)
  real(c_double) :: d
  integer(c_int) :: m
  call check(c_cfl1d_m3(d, safety, dtmin, m, gm, dmap(bx, ix), dmap(ro, ix), dmap(pr, ix), dmap(vx, ix), &
       dmap(vy, ix), dmap(by, ix), dmap(vz, ix), dmap(bz, ix), dmap(dx, ix), int(ix, c_int), c_null_ptr), 'cfl_m3')
  call dunmap(bx, 0); call dunmap(ro, 0); call dunmap(pr, 0); call dunmap(vx, 0); call dunmap(vy, 0)
  call dunmap(by, 0); call dunmap(vz, 0); call dunmap(bz, 0); call dunmap(dx, 0)
  dt = d
  merr = m
end subroutine cfl_m3

subroutine roe_m(ro, pr, vx, vy, by, bx, bxm, dt, gm, dx, ix)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix
  double precision, target :: ro(ix), pr(ix), vx(ix), vy(ix), by(ix)
  double precision, target, intent(in) :: bx(ix), bxm(ix), dx(ix)
  double precision, intent(in) :: dt, gm
  call check(c_roe1d(dmap(ro, ix), dmap(pr, ix), dmap(vx, ix), dmap(vy, ix), dmap(by, ix), dmap(bx, ix), &
       dmap(bxm, ix), dt, gm, dmap(dx, ix), int(ix, c_int), c_null_ptr), 'roe_m')
  call dunmap(ro, 1); call dunmap(pr, 1); call dunmap(vx, 1); call dunmap(vy, 1); call dunmap(by, 1)
  call dunmap(bx, 0); call dunmap(bxm, 0); call dunmap(dx, 0)
end subroutine roe_m

subroutine cfl_m(dt, safety, dtmin, merr, gm, bx, ro, pr, vx, vy, by, dx, ix)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix
  integer, intent(out) :: merr
  double precision, intent(out) :: dt
  double precision, intent(in) :: safety, dtmin, gm
  double precision, target, intent(in) :: bx(ix), ro(ix), pr(ix), vx(ix), vy(ix), by(ix), dx(ix)
  real(c_double) :: d
  integer(c_int) :: m
  call check(c_cfl1d(d, safety, dtmin, m, gm, dmap(bx, ix), dmap(ro, ix), dmap(pr, ix), dmap(vx, ix), &
       dmap(vy, ix), dmap(by, ix), dmap(dx, ix), int(ix, c_int), c_null_ptr), 'cfl_m')
  call dunmap(bx, 0); call dunmap(ro, 0); call dunmap(pr, 0); call dunmap(vx, 0)
  call dunmap(vy, 0); call dunmap(by, 0); call dunmap(dx, 0)
  dt = d
  merr = m
end subroutine cfl_m


subroutine bdfrex(mbnd, margin, qq, ix)
  use cans_gpu_iface
  integer mbnd, margin, ix
  double precision qq(ix)
  call bnd1(0, mbnd, margin, qq, 0.0d0, ix)
end subroutine bdfrex

subroutine bdsppx(mbnd, margin, qq, ix)
  use cans_gpu_iface
  integer mbnd, margin, ix
  double precision qq(ix)
  call bnd1(1, mbnd, margin, qq, 0.0d0, ix)
end subroutine bdsppx

subroutine bdspnx(mbnd, margin, qq, ix)
  use cans_gpu_iface
  integer mbnd, margin, ix
  double precision qq(ix)
  call bnd1(2, mbnd, margin, qq, 0.0d0, ix)
end subroutine bdspnx

subroutine bdsmpx(mbnd, margin, qq, ix)
  use cans_gpu_iface
  integer mbnd, margin, ix
  double precision qq(ix)
  call bnd1(3, mbnd, margin, qq, 0.0d0, ix)
end subroutine bdsmpx

subroutine bdsmnx(mbnd, margin, qq, ix)
  use cans_gpu_iface
  integer mbnd, margin, ix
  double precision qq(ix)
  call bnd1(4, mbnd, margin, qq, 0.0d0, ix)
end subroutine bdsmnx

subroutine bdcnsx(mbnd, margin, qq, q0, ix)
  use cans_gpu_iface
  integer mbnd, margin, ix
  double precision qq(ix), q0
  call bnd1(5, mbnd, margin, qq, q0, ix)
end subroutine bdcnsx

subroutine bdperx(margin0, margin1, qq, ix)
  use cans_gpu_iface
  integer margin0, margin1, ix
  double precision qq(ix)
  call bnd1(7, margin0, margin1, qq, 0.0d0, ix)
end subroutine bdperx

subroutine bdfrdx(mbnd, margin, qq, dxm, ix)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: mbnd, margin, ix
  double precision, target :: qq(ix)
  double precision, target, intent(in) :: dxm(ix)
  call check(c_bnd1d(BD_FRD, int(mbnd, c_int), int(margin, c_int), dmap(qq, ix), 0.0d0, dmap(dxm, ix), &
       int(ix, c_int), c_null_ptr), 'bdfrdx')
  call dunmap(qq, 1); call dunmap(dxm, 0)
end subroutine bdfrdx

!=======================================================================
! the helper routines the engines are built from, with the reference argument
! lists (device counterparts; the engines above do not call them)
!=======================================================================
! cans1d/hdmlw/mlwhalf.f:2
subroutine mlwhalf(u, un, du, dt, f, dxi, dxim, ix)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix
  double precision, target, intent(in) :: u(*)
  double precision, target :: un(*)
  double precision, target :: du(*)
  double precision, intent(in) :: dt
  double precision, target, intent(in) :: f(*)
  double precision, target, intent(in) :: dxi(*)
  double precision, target, intent(in) :: dxim(*)
  call check(c_mlwhalf1d(dmap(u, ix), dmap(un, ix), dmap(du, ix), dt, dmap(f, ix), dmap(dxi, ix), dmap(dxim, ix), &
       int(ix, c_int), c_null_ptr), 'mlwhalf')
  call dunmap(u, 0); call dunmap(un, 1); call dunmap(du, 1); call dunmap(f, 0)
  call dunmap(dxi, 0); call dunmap(dxim, 0)
end subroutine mlwhalf

! cans1d/hdmlw/mlwfull.f:2
subroutine mlwfull(du, dt, f, dxi, ix)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix
  double precision, target :: du(*)
  double precision, intent(in) :: dt
  double precision, target, intent(in) :: f(*)
  double precision, target, intent(in) :: dxi(*)
  call check(c_mlwfull1d(dmap(du, ix), dt, dmap(f, ix), dmap(dxi, ix), int(ix, c_int), c_null_ptr), 'mlwfull')
  call dunmap(du, 1); call dunmap(f, 0); call dunmap(dxi, 0)
end subroutine mlwfull

! cans1d/hdmlw/mlwartv.f:2
subroutine mlwartv(u, du, dt, qx, dxi, dxim, ix)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix
  double precision, target, intent(in) :: u(*)
  double precision, target :: du(*)
  double precision, intent(in) :: dt
  double precision, target, intent(in) :: qx(*)
  double precision, target, intent(in) :: dxi(*)
  double precision, target, intent(in) :: dxim(*)
  call check(c_mlwartv1d(dmap(u, ix), dmap(du, ix), dt, dmap(qx, ix), dmap(dxi, ix), dmap(dxim, ix), int(ix, c_int), &
       c_null_ptr), 'mlwartv')
  call dunmap(u, 0); call dunmap(du, 1); call dunmap(qx, 0); call dunmap(dxi, 0)
  call dunmap(dxim, 0)
end subroutine mlwartv

! cans1d/hdmlw/mlwsrch.f:2
subroutine mlwsrch(un, du, dt, s, ix)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix
  double precision, target :: un(*)
  double precision, target :: du(*)
  double precision, intent(in) :: dt
  double precision, target, intent(in) :: s(*)
  call check(c_mlwsrch1d(dmap(un, ix), dmap(du, ix), dt, dmap(s, ix), int(ix, c_int), c_null_ptr), 'mlwsrch')
  call dunmap(un, 1); call dunmap(du, 1); call dunmap(s, 0)
end subroutine mlwsrch

! cans1d/hdmlw/mlwsrcf.f:2
subroutine mlwsrcf(du, dt, s, ux0, ux1, ix)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix
  double precision, target :: du(*)
  double precision, intent(in) :: dt
  double precision, target, intent(in) :: s(*)
  double precision, target, intent(in) :: ux0(*)
  double precision, target, intent(in) :: ux1(*)
  call check(c_mlwsrcf1d(dmap(du, ix), dt, dmap(s, ix), dmap(ux0, ix), dmap(ux1, ix), int(ix, c_int), c_null_ptr), 'mlwsrcf')
  call dunmap(du, 1); call dunmap(s, 0); call dunmap(ux0, 0); call dunmap(ux1, 0)
end subroutine mlwsrcf

! cans1d/hdroe/roeflux_m.f:2-3
subroutine roeflux_m(fro, fee, frx, fry, fby, gm, row, prw, vxw, vyw, bxw, byw, ix)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix
  double precision, target :: fro(*)
  double precision, target :: fee(*)
  double precision, target :: frx(*)
  double precision, target :: fry(*)
  double precision, target :: fby(*)
  double precision, intent(in) :: gm
  double precision, target, intent(in) :: row(*)
  double precision, target, intent(in) :: prw(*)
  double precision, target, intent(in) :: vxw(*)
  double precision, target, intent(in) :: vyw(*)
  double precision, target, intent(in) :: bxw(*)
  double precision, target, intent(in) :: byw(*)
  call check(c_roeflux_m1d(dmap(fro, ix), dmap(fee, ix), dmap(frx, ix), dmap(fry, ix), dmap(fby, ix), gm, dmap(row, 2*ix), &
       dmap(prw, 2*ix), dmap(vxw, 2*ix), dmap(vyw, 2*ix), dmap(bxw, 2*ix), dmap(byw, 2*ix), int(ix, c_int), &
       c_null_ptr), 'roeflux_m')
  call dunmap(fro, 1); call dunmap(fee, 1); call dunmap(frx, 1); call dunmap(fry, 1)
  call dunmap(fby, 1); call dunmap(row, 0); call dunmap(prw, 0); call dunmap(vxw, 0)
  call dunmap(vyw, 0); call dunmap(bxw, 0); call dunmap(byw, 0)
end subroutine roeflux_m

! cans1d/hdroe/tvdminmod.f:2
subroutine tvdminmod(da, daw, ix)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix
  double precision, target, intent(in) :: da(*)
  double precision, target :: daw(*)
  call check(c_tvdminmod1d(dmap(da, ix), dmap(daw, 2*ix), int(ix, c_int), c_null_ptr), 'tvdminmod')
  call dunmap(da, 0); call dunmap(daw, 1)
end subroutine tvdminmod

