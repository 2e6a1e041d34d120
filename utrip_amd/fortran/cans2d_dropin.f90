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
