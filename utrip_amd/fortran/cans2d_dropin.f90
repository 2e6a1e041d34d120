!=======================================================================
! libcans2d_gpu -- the cans2d hot-path routines with the reference's names
! and argument lists (cans2d/hdmlw/mlw_m.f:2-3, cans2d/hdroe/roe_m.f:2,
! cans2d/common/cfl_m.f:2-3, cans2d/hdmlw/mlw_m3.f:2-3, cans2d/hdroe/roe_m3.f:2,
! cans2d/hdmlw/mlw_m_c.f:2-3,
! cans2d/common/cfl_m3.f:2-3, cans2d/bc/bd*{x,y}.f), computed on the GPU by
! libcans_gpu.so.
!=======================================================================
subroutine mlw_m(ro, pr, vx, vy, bx, by, az, dt, qav, gm, dx, dxm, ix, dy, dym, jx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx
  double precision, target :: ro(ix, jx), pr(ix, jx), vx(ix, jx), vy(ix, jx), bx(ix, jx), by(ix, jx), az(ix, jx)
  double precision, target, intent(in) :: dx(ix), dxm(ix), dy(jx), dym(jx)
  double precision, intent(in) :: dt, qav, gm
  integer :: n
  n = ix * jx
  call check(c_mlw2d(dmap(ro, n), dmap(pr, n), dmap(vx, n), dmap(vy, n), dmap(bx, n), dmap(by, n), dmap(az, n), &
       dt, qav, gm, dmap(dx, ix), dmap(dxm, ix), int(ix, c_int), dmap(dy, jx), dmap(dym, jx), int(jx, c_int), &
       c_null_ptr), 'mlw_m')
  call dunmap(ro, 1); call dunmap(pr, 1); call dunmap(vx, 1); call dunmap(vy, 1)
  call dunmap(bx, 1); call dunmap(by, 1); call dunmap(az, 1)
  call dunmap(dx, 0); call dunmap(dxm, 0); call dunmap(dy, 0); call dunmap(dym, 0)
end subroutine mlw_m

subroutine mlw_m3(ro, pr, vx, vy, vz, bx, by, bz, az, dt, qav, gm, dx, dxm, ix, dy, dym, jx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx
  double precision, target :: ro(ix, jx), pr(ix, jx), vx(ix, jx), vy(ix, jx), vz(ix, jx)
  double precision, target :: bx(ix, jx), by(ix, jx), bz(ix, jx), az(ix, jx)
  double precision, target, intent(in) :: dx(ix), dxm(ix), dy(jx), dym(jx)
  double precision, intent(in) :: dt, qav, gm
  integer :: n
  n = ix * jx
  call check(c_mlw2d_m3(dmap(ro, n), dmap(pr, n), dmap(vx, n), dmap(vy, n), dmap(vz, n), dmap(bx, n), dmap(by, n), &
       dmap(bz, n), dmap(az, n), dt, qav, gm, dmap(dx, ix), dmap(dxm, ix), int(ix, c_int), dmap(dy, jx), &
       dmap(dym, jx), int(jx, c_int), c_null_ptr), 'mlw_m3')
  call dunmap(ro, 1); call dunmap(pr, 1); call dunmap(vx, 1); call dunmap(vy, 1); call dunmap(vz, 1)
  call dunmap(bx, 1); call dunmap(by, 1); call dunmap(bz, 1); call dunmap(az, 1)
  call dunmap(dx, 0); call dunmap(dxm, 0); call dunmap(dy, 0); call dunmap(dym, 0)
end subroutine mlw_m3

subroutine mlw_m_c(ro, pr, vx, vz, bx, bz, ay, dt, qav, gm, x, xm, dx, dxm, ix, dz, dzm, jx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx
  double precision, target :: ro(ix, jx), pr(ix, jx), vx(ix, jx), vz(ix, jx), bx(ix, jx), bz(ix, jx), ay(ix, jx)
  double precision, target, intent(in) :: x(ix), xm(ix), dx(ix), dxm(ix), dz(jx), dzm(jx)
  double precision, intent(in) :: dt, qav, gm
  integer :: n
  n = ix * jx
  call check(c_mlw2d_c(dmap(ro, n), dmap(pr, n), dmap(vx, n), dmap(vz, n), dmap(bx, n), dmap(bz, n), dmap(ay, n), &
       dt, qav, gm, dmap(x, ix), dmap(xm, ix), dmap(dx, ix), dmap(dxm, ix), int(ix, c_int), dmap(dz, jx), &
       dmap(dzm, jx), int(jx, c_int), c_null_ptr), 'mlw_m_c')
  call dunmap(ro, 1); call dunmap(pr, 1); call dunmap(vx, 1); call dunmap(vz, 1)
  call dunmap(bx, 1); call dunmap(bz, 1); call dunmap(ay, 1)
  call dunmap(x, 0); call dunmap(xm, 0); call dunmap(dx, 0); call dunmap(dxm, 0); call dunmap(dz, 0); call dunmap(dzm, 0)
end subroutine mlw_m_c

! cans2d/hdmlw/mlw_m_s.f:1-2: spherical (r, theta) with Az
subroutine mlw_m_s(ro, pr, vx, vy, bx, by, az, dt, qav, gm, x, xm, dx, dxm, ix, y, ym, dy, dym, jx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx
  double precision, target :: ro(ix, jx), pr(ix, jx), vx(ix, jx), vy(ix, jx), bx(ix, jx), by(ix, jx), az(ix, jx)
  double precision, target, intent(in) :: x(ix), xm(ix), dx(ix), dxm(ix), y(jx), ym(jx), dy(jx), dym(jx)
  double precision, intent(in) :: dt, qav, gm
  integer :: n
  n = ix * jx
  call check(c_mlw2d_s(dmap(ro, n), dmap(pr, n), dmap(vx, n), dmap(vy, n), dmap(bx, n), dmap(by, n), dmap(az, n), &
       dt, qav, gm, dmap(x, ix), dmap(xm, ix), dmap(dx, ix), dmap(dxm, ix), int(ix, c_int), dmap(y, jx), &
       dmap(ym, jx), dmap(dy, jx), dmap(dym, jx), int(jx, c_int), c_null_ptr), 'mlw_m_s')
  call dunmap(ro, 1); call dunmap(pr, 1); call dunmap(vx, 1); call dunmap(vy, 1)
  call dunmap(bx, 1); call dunmap(by, 1); call dunmap(az, 1)
  call dunmap(x, 0); call dunmap(xm, 0); call dunmap(dx, 0); call dunmap(dxm, 0)
  call dunmap(y, 0); call dunmap(ym, 0); call dunmap(dy, 0); call dunmap(dym, 0)
end subroutine mlw_m_s

! cans2d/common/cfl_m_s.f:1-2
subroutine cfl_m_s(dt, safety, dtmin, merr, gm, ro, pr, vx, vy, bx, by, x, dx, ix, dy, jx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx
  integer, intent(out) :: merr
  double precision, intent(out) :: dt
  double precision, intent(in) :: safety, dtmin, gm
  double precision, target, intent(in) :: ro(ix, jx), pr(ix, jx), vx(ix, jx), vy(ix, jx), bx(ix, jx), by(ix, jx)
  double precision, target, intent(in) :: x(ix), dx(ix), dy(jx)
  real(c_double) :: d
  integer(c_int) :: m
  integer :: n
  n = ix * jx
  call check(c_cfl2d_s(d, safety, dtmin, m, gm, dmap(ro, n), dmap(pr, n), dmap(vx, n), dmap(vy, n), dmap(bx, n), &
       dmap(by, n), dmap(x, ix), dmap(dx, ix), int(ix, c_int), dmap(dy, jx), int(jx, c_int), c_null_ptr), 'cfl_m_s')
  call dunmap(ro, 0); call dunmap(pr, 0); call dunmap(vx, 0); call dunmap(vy, 0)
  call dunmap(bx, 0); call dunmap(by, 0); call dunmap(x, 0); call dunmap(dx, 0); call dunmap(dy, 0)
  dt = d
  merr = m
end subroutine cfl_m_s

subroutine roe_m3(ro, pr, vx, vy, vz, bx, by, bz, az, dt, gm, dx, ix, dy, jx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx
  double precision, target :: ro(ix, jx), pr(ix, jx), vx(ix, jx), vy(ix, jx), vz(ix, jx)
  double precision, target :: bx(ix, jx), by(ix, jx), bz(ix, jx), az(ix, jx)
  double precision, target, intent(in) :: dx(ix), dy(jx)
  double precision, intent(in) :: dt, gm
  integer :: n
  n = ix * jx
  call check(c_roe2d_m3(dmap(ro, n), dmap(pr, n), dmap(vx, n), dmap(vy, n), dmap(vz, n), dmap(bx, n), dmap(by, n), &
       dmap(bz, n), dmap(az, n), dt, gm, dmap(dx, ix), int(ix, c_int), dmap(dy, jx), int(jx, c_int), c_null_ptr), &
       'roe_m3')
  call dunmap(ro, 1); call dunmap(pr, 1); call dunmap(vx, 1); call dunmap(vy, 1); call dunmap(vz, 1)
  call dunmap(bx, 1); call dunmap(by, 1); call dunmap(bz, 1); call dunmap(az, 1)
  call dunmap(dx, 0); call dunmap(dy, 0)
end subroutine roe_m3

subroutine cfl_m3(dt, safety, dtmin, merr, gm, ro, pr, vx, vy, vz, bx, by, bz, dx, ix, dy, jx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx
  integer, intent(out) :: merr
  double precision, intent(out) :: dt
  double precision, intent(in) :: safety, dtmin, gm
  double precision, target, intent(in) :: ro(ix, jx), pr(ix, jx), vx(ix, jx), vy(ix, jx), vz(ix, jx)
  double precision, target, intent(in) :: bx(ix, jx), by(ix, jx), bz(ix, jx), dx(ix), dy(jx)
  real(c_double) :: d
  integer(c_int) :: m
  integer :: n
  n = ix * jx
  call check(c_cfl2d_m3(d, safety, dtmin, m, gm, dmap(ro, n), dmap(pr, n), dmap(vx, n), dmap(vy, n), dmap(vz, n), &
       dmap(bx, n), dmap(by, n), dmap(bz, n), dmap(dx, ix), int(ix, c_int), dmap(dy, jx), int(jx, c_int), &
       c_null_ptr), 'cfl_m3')
  call dunmap(ro, 0); call dunmap(pr, 0); call dunmap(vx, 0); call dunmap(vy, 0); call dunmap(vz, 0)
  call dunmap(bx, 0); call dunmap(by, 0); call dunmap(bz, 0); call dunmap(dx, 0); call dunmap(dy, 0)
  dt = d
  merr = m
end subroutine cfl_m3

subroutine roe_m(ro, pr, vx, vy, bx, by, az, dt, gm, dx, ix, dy, jx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx
  double precision, target :: ro(ix, jx), pr(ix, jx), vx(ix, jx), vy(ix, jx), bx(ix, jx), by(ix, jx), az(ix, jx)
  double precision, target, intent(in) :: dx(ix), dy(jx)
  double precision, intent(in) :: dt, gm
  integer :: n
  n = ix * jx
  call check(c_roe2d(dmap(ro, n), dmap(pr, n), dmap(vx, n), dmap(vy, n), dmap(bx, n), dmap(by, n), dmap(az, n), &
       dt, gm, dmap(dx, ix), int(ix, c_int), dmap(dy, jx), int(jx, c_int), c_null_ptr), 'roe_m')
  call dunmap(ro, 1); call dunmap(pr, 1); call dunmap(vx, 1); call dunmap(vy, 1)
  call dunmap(bx, 1); call dunmap(by, 1); call dunmap(az, 1)
  call dunmap(dx, 0); call dunmap(dy, 0)
end subroutine roe_m

subroutine cfl_m(dt, safety, dtmin, merr, gm, ro, pr, vx, vy, bx, by, dx, ix, dy, jx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx
  integer, intent(out) :: merr
  double precision, intent(out) :: dt
  double precision, intent(in) :: safety, dtmin, gm
  double precision, target, intent(in) :: ro(ix, jx), pr(ix, jx), vx(ix, jx), vy(ix, jx), bx(ix, jx), by(ix, jx)
  double precision, target, intent(in) :: dx(ix), dy(jx)
  real(c_double) :: d
  integer(c_int) :: m
  integer :: n
  n = ix * jx
  call check(c_cfl2d(d, safety, dtmin, m, gm, dmap(ro, n), dmap(pr, n), dmap(vx, n), dmap(vy, n), dmap(bx, n), &
       dmap(by, n), dmap(dx, ix), int(ix, c_int), dmap(dy, jx), int(jx, c_int), c_null_ptr), 'cfl_m')
  call dunmap(ro, 0); call dunmap(pr, 0); call dunmap(vx, 0); call dunmap(vy, 0)
  call dunmap(bx, 0); call dunmap(by, 0); call dunmap(dx, 0); call dunmap(dy, 0)
  dt = d
  merr = m
end subroutine cfl_m



subroutine bdfrex(mbnd, margin, qq, ix, jx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx
  double precision qq(ix, jx)
  call bnd2(0, 0, mbnd, margin, qq, 0.0d0, ix, jx)
end subroutine bdfrex

subroutine bdsppx(mbnd, margin, qq, ix, jx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx
  double precision qq(ix, jx)
  call bnd2(0, 1, mbnd, margin, qq, 0.0d0, ix, jx)
end subroutine bdsppx

subroutine bdspnx(mbnd, margin, qq, ix, jx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx
  double precision qq(ix, jx)
  call bnd2(0, 2, mbnd, margin, qq, 0.0d0, ix, jx)
end subroutine bdspnx

subroutine bdsmpx(mbnd, margin, qq, ix, jx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx
  double precision qq(ix, jx)
  call bnd2(0, 3, mbnd, margin, qq, 0.0d0, ix, jx)
end subroutine bdsmpx

subroutine bdsmnx(mbnd, margin, qq, ix, jx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx
  double precision qq(ix, jx)
  call bnd2(0, 4, mbnd, margin, qq, 0.0d0, ix, jx)
end subroutine bdsmnx

subroutine bdcnsx(mbnd, margin, qq, q0, ix, jx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx
  double precision qq(ix, jx), q0
  call bnd2(0, 5, mbnd, margin, qq, q0, ix, jx)
end subroutine bdcnsx

subroutine bdperx(margin0, margin1, qq, ix, jx)
  use cans_gpu_iface
  integer margin0, margin1, ix, jx
  double precision qq(ix, jx)
  call bnd2(0, 7, margin0, margin1, qq, 0.0d0, ix, jx)
end subroutine bdperx

subroutine bdfrdx(mbnd, margin, qq, dxm, ix, jx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx
  double precision qq(ix, jx), dxm(*)
  call bnd2d_frd(0, mbnd, margin, qq, dxm, ix, ix, jx)
end subroutine bdfrdx

subroutine bdfrey(mbnd, margin, qq, ix, jx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx
  double precision qq(ix, jx)
  call bnd2(1, 0, mbnd, margin, qq, 0.0d0, ix, jx)
end subroutine bdfrey

subroutine bdsppy(mbnd, margin, qq, ix, jx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx
  double precision qq(ix, jx)
  call bnd2(1, 1, mbnd, margin, qq, 0.0d0, ix, jx)
end subroutine bdsppy

subroutine bdspny(mbnd, margin, qq, ix, jx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx
  double precision qq(ix, jx)
  call bnd2(1, 2, mbnd, margin, qq, 0.0d0, ix, jx)
end subroutine bdspny

subroutine bdsmpy(mbnd, margin, qq, ix, jx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx
  double precision qq(ix, jx)
  call bnd2(1, 3, mbnd, margin, qq, 0.0d0, ix, jx)
end subroutine bdsmpy

subroutine bdsmny(mbnd, margin, qq, ix, jx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx
  double precision qq(ix, jx)
  call bnd2(1, 4, mbnd, margin, qq, 0.0d0, ix, jx)
end subroutine bdsmny

subroutine bdcnsy(mbnd, margin, qq, q0, ix, jx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx
  double precision qq(ix, jx), q0
  call bnd2(1, 5, mbnd, margin, qq, q0, ix, jx)
end subroutine bdcnsy

subroutine bdpery(margin0, margin1, qq, ix, jx)
  use cans_gpu_iface
  integer margin0, margin1, ix, jx
  double precision qq(ix, jx)
  call bnd2(1, 7, margin0, margin1, qq, 0.0d0, ix, jx)
end subroutine bdpery

subroutine bdfrdy(mbnd, margin, qq, dym, ix, jx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx
  double precision qq(ix, jx), dym(*)
  call bnd2d_frd(1, mbnd, margin, qq, dym, jx, ix, jx)
end subroutine bdfrdy

!=======================================================================
! the helper routines the engines are built from, with the reference argument
! lists (device counterparts; the engines above do not call them)
!=======================================================================
! cans2d/hdmlw/mlwhalf.f:2
subroutine mlwhalf(u, un, du, dt, fx, dxi, dxim, ix, fy, dyi, dyim, jx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx
  double precision, target, intent(in) :: u(*)
  double precision, target :: un(*)
  double precision, target :: du(*)
  double precision, intent(in) :: dt
  double precision, target, intent(in) :: fx(*)
  double precision, target, intent(in) :: dxi(*)
  double precision, target, intent(in) :: dxim(*)
  double precision, target, intent(in) :: fy(*)
  double precision, target, intent(in) :: dyi(*)
  double precision, target, intent(in) :: dyim(*)
  integer :: n
  n = ix * jx
  call check(c_mlwhalf2d(dmap(u, n), dmap(un, n), dmap(du, n), dt, dmap(fx, n), dmap(dxi, ix), dmap(dxim, ix), &
       int(ix, c_int), dmap(fy, n), dmap(dyi, jx), dmap(dyim, jx), int(jx, c_int), c_null_ptr), 'mlwhalf')
  call dunmap(u, 0); call dunmap(un, 1); call dunmap(du, 1); call dunmap(fx, 0)
  call dunmap(dxi, 0); call dunmap(dxim, 0); call dunmap(fy, 0); call dunmap(dyi, 0)
  call dunmap(dyim, 0)
end subroutine mlwhalf

! cans2d/hdmlw/mlwfull.f:2
subroutine mlwfull(du, dt, fx, dxi, ux0, ux1, ix, fy, dyi, uy0, uy1, jx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx
  double precision, target :: du(*)
  double precision, intent(in) :: dt
  double precision, target, intent(in) :: fx(*)
  double precision, target, intent(in) :: dxi(*)
  double precision, target, intent(in) :: ux0(*)
  double precision, target, intent(in) :: ux1(*)
  double precision, target, intent(in) :: fy(*)
  double precision, target, intent(in) :: dyi(*)
  double precision, target, intent(in) :: uy0(*)
  double precision, target, intent(in) :: uy1(*)
  integer :: n
  n = ix * jx
  call check(c_mlwfull2d(dmap(du, n), dt, dmap(fx, n), dmap(dxi, ix), dmap(ux0, ix), dmap(ux1, ix), int(ix, c_int), &
       dmap(fy, n), dmap(dyi, jx), dmap(uy0, jx), dmap(uy1, jx), int(jx, c_int), c_null_ptr), 'mlwfull')
  call dunmap(du, 1); call dunmap(fx, 0); call dunmap(dxi, 0); call dunmap(ux0, 0)
  call dunmap(ux1, 0); call dunmap(fy, 0); call dunmap(dyi, 0); call dunmap(uy0, 0)
  call dunmap(uy1, 0)
end subroutine mlwfull

! cans2d/hdmlw/mlwartv.f:2
subroutine mlwartv(u, du, dt, qx, dxi, dxim, ix, qy, dyi, dyim, jx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx
  double precision, target, intent(in) :: u(*)
  double precision, target :: du(*)
  double precision, intent(in) :: dt
  double precision, target, intent(in) :: qx(*)
  double precision, target, intent(in) :: dxi(*)
  double precision, target, intent(in) :: dxim(*)
  double precision, target, intent(in) :: qy(*)
  double precision, target, intent(in) :: dyi(*)
  double precision, target, intent(in) :: dyim(*)
  integer :: n
  n = ix * jx
  call check(c_mlwartv2d(dmap(u, n), dmap(du, n), dt, dmap(qx, n), dmap(dxi, ix), dmap(dxim, ix), int(ix, c_int), &
       dmap(qy, n), dmap(dyi, jx), dmap(dyim, jx), int(jx, c_int), c_null_ptr), 'mlwartv')
  call dunmap(u, 0); call dunmap(du, 1); call dunmap(qx, 0); call dunmap(dxi, 0)
  call dunmap(dxim, 0); call dunmap(qy, 0); call dunmap(dyi, 0); call dunmap(dyim, 0)
end subroutine mlwartv

! cans2d/hdmlw/mlwsrch.f:2
subroutine mlwsrch(un, du, dt, s, ix, jx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx
  double precision, target :: un(*)
  double precision, target :: du(*)
  double precision, intent(in) :: dt
  double precision, target, intent(in) :: s(*)
  integer :: n
  n = ix * jx
  call check(c_mlwsrch2d(dmap(un, n), dmap(du, n), dt, dmap(s, n), int(ix, c_int), int(jx, c_int), c_null_ptr), 'mlwsrch')
  call dunmap(un, 1); call dunmap(du, 1); call dunmap(s, 0)
end subroutine mlwsrch

! cans2d/hdmlw/mlwsrcf.f:2
subroutine mlwsrcf(du, dt, s, ux0, ux1, ix, uy0, uy1, jx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx
  double precision, target :: du(*)
  double precision, intent(in) :: dt
  double precision, target, intent(in) :: s(*)
  double precision, target, intent(in) :: ux0(*)
  double precision, target, intent(in) :: ux1(*)
  double precision, target, intent(in) :: uy0(*)
  double precision, target, intent(in) :: uy1(*)
  integer :: n
  n = ix * jx
  call check(c_mlwsrcf2d(dmap(du, n), dt, dmap(s, n), dmap(ux0, ix), dmap(ux1, ix), int(ix, c_int), dmap(uy0, jx), &
       dmap(uy1, jx), int(jx, c_int), c_null_ptr), 'mlwsrcf')
  call dunmap(du, 1); call dunmap(s, 0); call dunmap(ux0, 0); call dunmap(ux1, 0)
  call dunmap(uy0, 0); call dunmap(uy1, 0)
end subroutine mlwsrcf

! cans2d/hdroe/roeflux_m.f:1-2
subroutine roeflux_m(fro, fee, frx, fry, fby, gm, row, prw, vxw, vyw, bxw, byw, ix, jx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx
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
  integer :: n
  n = ix * jx
  call check(c_roeflux_m2d(dmap(fro, n), dmap(fee, n), dmap(frx, n), dmap(fry, n), dmap(fby, n), gm, dmap(row, 2*n), &
       dmap(prw, 2*n), dmap(vxw, 2*n), dmap(vyw, 2*n), dmap(bxw, 2*n), dmap(byw, 2*n), int(ix, c_int), &
       int(jx, c_int), c_null_ptr), 'roeflux_m')
  call dunmap(fro, 1); call dunmap(fee, 1); call dunmap(frx, 1); call dunmap(fry, 1)
  call dunmap(fby, 1); call dunmap(row, 0); call dunmap(prw, 0); call dunmap(vxw, 0)
  call dunmap(vyw, 0); call dunmap(bxw, 0); call dunmap(byw, 0)
end subroutine roeflux_m

! cans2d/hdroe/tvdminmod.f:1
subroutine tvdminmod(mdir, da, daw, ix, jx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: mdir, ix, jx
  double precision, target, intent(in) :: da(*)
  double precision, target :: daw(*)
  integer :: n
  n = ix * jx
  call check(c_tvdminmod2d(int(mdir, c_int), dmap(da, n), dmap(daw, 2*n), int(ix, c_int), int(jx, c_int), c_null_ptr), 'tvdminmod')
  call dunmap(da, 0); call dunmap(daw, 1)
end subroutine tvdminmod

