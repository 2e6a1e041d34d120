!=======================================================================
! libcans3d_gpu -- the cans3d hot-path routines with the reference's names
! and argument lists (cans3d/hdmlw/mlw_m.f:2-3, cans3d/hdmlw/mlw_m_c.f:2-3, cans3d/hdmlw/mlw_m_s.f:2-3,
! cans3d/common/cfl_m_s.f:1-2, cans3d/hdroe/roe_m.f:2,
! cans3d/common/cfl_m.f:2-3, cans3d/bc/bd*{x,y,z}.f) and the HLLD driver
! engine of cansxx-nosupport/md_mhdshktb/hdhlld-x/engine.f:2, computed on
! the GPU by libcans_gpu.so.
!=======================================================================
subroutine mlw_m(ro, pr, vx, vy, vz, bx, by, bz, dt, qav, gm, dx, dxm, ix, dy, dym, jx, dz, dzm, kx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx, kx
  double precision, target :: ro(*), pr(*), vx(*), vy(*), vz(*), bx(*), by(*), bz(*)
  double precision, target, intent(in) :: dx(ix), dxm(ix), dy(jx), dym(jx), dz(kx), dzm(kx)
  double precision, intent(in) :: dt, qav, gm
  integer :: n
  n = ix * jx * kx
  call check(c_mlw3d(dmap(ro, n), dmap(pr, n), dmap(vx, n), dmap(vy, n), dmap(vz, n), dmap(bx, n), dmap(by, n), &
       dmap(bz, n), dt, qav, gm, dmap(dx, ix), dmap(dxm, ix), int(ix, c_int), dmap(dy, jx), dmap(dym, jx), &
       int(jx, c_int), dmap(dz, kx), dmap(dzm, kx), int(kx, c_int), c_null_ptr), 'mlw_m')
  call unmap8(ro, pr, vx, vy, vz, bx, by, bz)
  call dunmap(dx, 0); call dunmap(dxm, 0); call dunmap(dy, 0); call dunmap(dym, 0)
  call dunmap(dz, 0); call dunmap(dzm, 0)
end subroutine mlw_m


! cans3d/hdmlw/mlw_m_c.f:2-3: cylindrical (r, phi, z), cell and face radii x, xm
subroutine mlw_m_c(ro, pr, vx, vy, vz, bx, by, bz, dt, qav, gm, x, xm, dx, dxm, ix, dy, dym, jx, dz, dzm, kx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx, kx
  double precision, target :: ro(*), pr(*), vx(*), vy(*), vz(*), bx(*), by(*), bz(*)
  double precision, target, intent(in) :: x(ix), xm(ix), dx(ix), dxm(ix), dy(jx), dym(jx), dz(kx), dzm(kx)
  double precision, intent(in) :: dt, qav, gm
  integer :: n
  n = ix * jx * kx
  call check(c_mlw3d_c(dmap(ro, n), dmap(pr, n), dmap(vx, n), dmap(vy, n), dmap(vz, n), dmap(bx, n), dmap(by, n), &
       dmap(bz, n), dt, qav, gm, dmap(x, ix), dmap(xm, ix), dmap(dx, ix), dmap(dxm, ix), int(ix, c_int), &
       dmap(dy, jx), dmap(dym, jx), int(jx, c_int), dmap(dz, kx), dmap(dzm, kx), int(kx, c_int), c_null_ptr), &
       'mlw_m_c')
  call unmap8(ro, pr, vx, vy, vz, bx, by, bz)
  call dunmap(x, 0); call dunmap(xm, 0)
  call dunmap(dx, 0); call dunmap(dxm, 0); call dunmap(dy, 0); call dunmap(dym, 0)
  call dunmap(dz, 0); call dunmap(dzm, 0)
end subroutine mlw_m_c


subroutine roe_m(ro, pr, vx, vy, vz, bx, by, bz, dt, gm, dx, ix, dy, jx, dz, kx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx, kx
  double precision, target :: ro(*), pr(*), vx(*), vy(*), vz(*), bx(*), by(*), bz(*)
  double precision, target, intent(in) :: dx(ix), dy(jx), dz(kx)
  double precision, intent(in) :: dt, gm
  integer :: n
  n = ix * jx * kx
  call check(c_roe3d(dmap(ro, n), dmap(pr, n), dmap(vx, n), dmap(vy, n), dmap(vz, n), dmap(bx, n), dmap(by, n), &
       dmap(bz, n), dt, gm, dmap(dx, ix), int(ix, c_int), dmap(dy, jx), int(jx, c_int), dmap(dz, kx), &
       int(kx, c_int), c_null_ptr), 'roe_m')
  call unmap8(ro, pr, vx, vy, vz, bx, by, bz)
  call dunmap(dx, 0); call dunmap(dy, 0); call dunmap(dz, 0)
end subroutine roe_m

subroutine cfl_m(dt, safety, dtmin, merr, gm, ro, pr, vx, vy, vz, bx, by, bz, dx, ix, dy, jx, dz, kx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx, kx
  integer, intent(out) :: merr
  double precision, intent(out) :: dt
  double precision, intent(in) :: safety, dtmin, gm
  double precision, target, intent(in) :: ro(*), pr(*), vx(*), vy(*), vz(*), bx(*), by(*), bz(*)
  double precision, target, intent(in) :: dx(ix), dy(jx), dz(kx)
  real(c_double) :: d
  integer(c_int) :: m
  integer :: n
  n = ix * jx * kx
  call check(c_cfl3d(d, safety, dtmin, m, gm, dmap(ro, n), dmap(pr, n), dmap(vx, n), dmap(vy, n), dmap(vz, n), &
       dmap(bx, n), dmap(by, n), dmap(bz, n), dmap(dx, ix), int(ix, c_int), dmap(dy, jx), int(jx, c_int), &
       dmap(dz, kx), int(kx, c_int), c_null_ptr), 'cfl_m')
  call dunmap(ro, 0); call dunmap(pr, 0); call dunmap(vx, 0); call dunmap(vy, 0)
  call dunmap(vz, 0); call dunmap(bx, 0); call dunmap(by, 0); call dunmap(bz, 0)
  call dunmap(dx, 0); call dunmap(dy, 0); call dunmap(dz, 0)
  dt = d
  merr = m
end subroutine cfl_m

! cans3d/hdmlw/mlw_m_s.f:2-3: spherical (r, theta, phi), radii x, xm and colatitudes y, ym
subroutine mlw_m_s(ro, pr, vx, vy, vz, bx, by, bz, dt, qav, gm, x, xm, dx, dxm, ix, y, ym, dy, dym, jx, &
     dz, dzm, kx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx, kx
  double precision, target :: ro(*), pr(*), vx(*), vy(*), vz(*), bx(*), by(*), bz(*)
  double precision, target, intent(in) :: x(ix), xm(ix), dx(ix), dxm(ix), y(jx), ym(jx), dy(jx), dym(jx)
  double precision, target, intent(in) :: dz(kx), dzm(kx)
  double precision, intent(in) :: dt, qav, gm
  integer :: n
  n = ix * jx * kx
  call check(c_mlw3d_s(dmap(ro, n), dmap(pr, n), dmap(vx, n), dmap(vy, n), dmap(vz, n), dmap(bx, n), dmap(by, n), &
       dmap(bz, n), dt, qav, gm, dmap(x, ix), dmap(xm, ix), dmap(dx, ix), dmap(dxm, ix), int(ix, c_int), &
       dmap(y, jx), dmap(ym, jx), dmap(dy, jx), dmap(dym, jx), int(jx, c_int), dmap(dz, kx), dmap(dzm, kx), &
       int(kx, c_int), c_null_ptr), 'mlw_m_s')
  call unmap8(ro, pr, vx, vy, vz, bx, by, bz)
  call dunmap(x, 0); call dunmap(xm, 0); call dunmap(dx, 0); call dunmap(dxm, 0)
  call dunmap(y, 0); call dunmap(ym, 0); call dunmap(dy, 0); call dunmap(dym, 0)
  call dunmap(dz, 0); call dunmap(dzm, 0)
end subroutine mlw_m_s

! cans3d/common/cfl_m_s.f:1-2
subroutine cfl_m_s(dt, safety, dtmin, merr, gm, ro, pr, vx, vy, vz, bx, by, bz, x, dx, ix, y, dy, jx, dz, kx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx, kx
  integer, intent(out) :: merr
  double precision, intent(out) :: dt
  double precision, intent(in) :: safety, dtmin, gm
  double precision, target, intent(in) :: ro(*), pr(*), vx(*), vy(*), vz(*), bx(*), by(*), bz(*)
  double precision, target, intent(in) :: x(ix), dx(ix), y(jx), dy(jx), dz(kx)
  real(c_double) :: d
  integer(c_int) :: m
  integer :: n
  n = ix * jx * kx
  call check(c_cfl3d_s(d, safety, dtmin, m, gm, dmap(ro, n), dmap(pr, n), dmap(vx, n), dmap(vy, n), &
       dmap(vz, n), dmap(bx, n), dmap(by, n), dmap(bz, n), dmap(x, ix), dmap(dx, ix), int(ix, c_int), &
       dmap(y, jx), dmap(dy, jx), int(jx, c_int), dmap(dz, kx), int(kx, c_int), c_null_ptr), 'cfl_m_s')
  call dunmap(ro, 0); call dunmap(pr, 0); call dunmap(vx, 0); call dunmap(vy, 0)
  call dunmap(vz, 0); call dunmap(bx, 0); call dunmap(by, 0); call dunmap(bz, 0)
  call dunmap(x, 0); call dunmap(dx, 0); call dunmap(y, 0); call dunmap(dy, 0); call dunmap(dz, 0)
  dt = d
  merr = m
end subroutine cfl_m_s

subroutine engine(ro, pr, vx, vy, vz, bx, by, bz, gm, mu, dt, dx, dxm, dy, dym, dz, dzm, ix, jx, kx, mfdim)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx, kx
  integer, target, intent(in) :: mfdim(3)
  double precision, target :: ro(*), pr(*), vx(*), vy(*), vz(*), bx(*), by(*), bz(*)
  double precision, target, intent(in) :: dx(ix), dxm(ix), dy(jx), dym(jx), dz(kx), dzm(kx)
  double precision, intent(in) :: gm, mu, dt
  integer :: n
  n = ix * jx * kx
  call check(c_hlld(dmap(ro, n), dmap(pr, n), dmap(vx, n), dmap(vy, n), dmap(vz, n), dmap(bx, n), dmap(by, n), &
       dmap(bz, n), gm, mu, dt, dmap(dx, ix), dmap(dy, jx), dmap(dz, kx), int(ix, c_int), int(jx, c_int), &
       int(kx, c_int), c_loc(mfdim(1)), c_null_ptr), 'engine')
  call unmap8(ro, pr, vx, vy, vz, bx, by, bz)
  call dunmap(dx, 0); call dunmap(dy, 0); call dunmap(dz, 0)
end subroutine engine



subroutine bdfrex(mbnd, margin, qq, ix, jx, kx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx, kx
  double precision qq(*)
  call bnd3(0, 0, mbnd, margin, qq, 0.0d0, ix, jx, kx)
end subroutine bdfrex

subroutine bdsppx(mbnd, margin, qq, ix, jx, kx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx, kx
  double precision qq(*)
  call bnd3(0, 1, mbnd, margin, qq, 0.0d0, ix, jx, kx)
end subroutine bdsppx

subroutine bdspnx(mbnd, margin, qq, ix, jx, kx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx, kx
  double precision qq(*)
  call bnd3(0, 2, mbnd, margin, qq, 0.0d0, ix, jx, kx)
end subroutine bdspnx

subroutine bdsmpx(mbnd, margin, qq, ix, jx, kx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx, kx
  double precision qq(*)
  call bnd3(0, 3, mbnd, margin, qq, 0.0d0, ix, jx, kx)
end subroutine bdsmpx

subroutine bdsmnx(mbnd, margin, qq, ix, jx, kx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx, kx
  double precision qq(*)
  call bnd3(0, 4, mbnd, margin, qq, 0.0d0, ix, jx, kx)
end subroutine bdsmnx

subroutine bdcnsx(mbnd, margin, qq, q0, ix, jx, kx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx, kx
  double precision qq(*), q0
  call bnd3(0, 5, mbnd, margin, qq, q0, ix, jx, kx)
end subroutine bdcnsx

subroutine bdperx(margin0, margin1, qq, ix, jx, kx)
  use cans_gpu_iface
  integer margin0, margin1, ix, jx, kx
  double precision qq(*)
  call bnd3(0, 7, margin0, margin1, qq, 0.0d0, ix, jx, kx)
end subroutine bdperx

subroutine bdfrdx(mbnd, margin, qq, dm, ix, jx, kx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx, kx
  double precision qq(*), dm(*)
  call bnd3_frd(0, mbnd, margin, qq, dm, ix, ix, jx, kx)
end subroutine bdfrdx

subroutine bdfrey(mbnd, margin, qq, ix, jx, kx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx, kx
  double precision qq(*)
  call bnd3(1, 0, mbnd, margin, qq, 0.0d0, ix, jx, kx)
end subroutine bdfrey

subroutine bdsppy(mbnd, margin, qq, ix, jx, kx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx, kx
  double precision qq(*)
  call bnd3(1, 1, mbnd, margin, qq, 0.0d0, ix, jx, kx)
end subroutine bdsppy

subroutine bdspny(mbnd, margin, qq, ix, jx, kx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx, kx
  double precision qq(*)
  call bnd3(1, 2, mbnd, margin, qq, 0.0d0, ix, jx, kx)
end subroutine bdspny

subroutine bdsmpy(mbnd, margin, qq, ix, jx, kx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx, kx
  double precision qq(*)
  call bnd3(1, 3, mbnd, margin, qq, 0.0d0, ix, jx, kx)
end subroutine bdsmpy

subroutine bdsmny(mbnd, margin, qq, ix, jx, kx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx, kx
  double precision qq(*)
  call bnd3(1, 4, mbnd, margin, qq, 0.0d0, ix, jx, kx)
end subroutine bdsmny

subroutine bdcnsy(mbnd, margin, qq, q0, ix, jx, kx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx, kx
  double precision qq(*), q0
  call bnd3(1, 5, mbnd, margin, qq, q0, ix, jx, kx)
end subroutine bdcnsy

subroutine bdpery(margin0, margin1, qq, ix, jx, kx)
  use cans_gpu_iface
  integer margin0, margin1, ix, jx, kx
  double precision qq(*)
  call bnd3(1, 7, margin0, margin1, qq, 0.0d0, ix, jx, kx)
end subroutine bdpery

subroutine bdfrdy(mbnd, margin, qq, dm, ix, jx, kx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx, kx
  double precision qq(*), dm(*)
  call bnd3_frd(1, mbnd, margin, qq, dm, jx, ix, jx, kx)
end subroutine bdfrdy

subroutine bdfrez(mbnd, margin, qq, ix, jx, kx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx, kx
  double precision qq(*)
  call bnd3(2, 0, mbnd, margin, qq, 0.0d0, ix, jx, kx)
end subroutine bdfrez

subroutine bdsppz(mbnd, margin, qq, ix, jx, kx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx, kx
  double precision qq(*)
  call bnd3(2, 1, mbnd, margin, qq, 0.0d0, ix, jx, kx)
end subroutine bdsppz

subroutine bdspnz(mbnd, margin, qq, ix, jx, kx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx, kx
  double precision qq(*)
  call bnd3(2, 2, mbnd, margin, qq, 0.0d0, ix, jx, kx)
end subroutine bdspnz

subroutine bdsmpz(mbnd, margin, qq, ix, jx, kx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx, kx
  double precision qq(*)
  call bnd3(2, 3, mbnd, margin, qq, 0.0d0, ix, jx, kx)
end subroutine bdsmpz

subroutine bdsmnz(mbnd, margin, qq, ix, jx, kx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx, kx
  double precision qq(*)
  call bnd3(2, 4, mbnd, margin, qq, 0.0d0, ix, jx, kx)
end subroutine bdsmnz

subroutine bdcnsz(mbnd, margin, qq, q0, ix, jx, kx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx, kx
  double precision qq(*), q0
  call bnd3(2, 5, mbnd, margin, qq, q0, ix, jx, kx)
end subroutine bdcnsz

subroutine bdperz(margin0, margin1, qq, ix, jx, kx)
  use cans_gpu_iface
  integer margin0, margin1, ix, jx, kx
  double precision qq(*)
  call bnd3(2, 7, margin0, margin1, qq, 0.0d0, ix, jx, kx)
end subroutine bdperz

subroutine bdfrdz(mbnd, margin, qq, dm, ix, jx, kx)
  use cans_gpu_iface
  integer mbnd, margin, ix, jx, kx
  double precision qq(*), dm(*)
  call bnd3_frd(2, mbnd, margin, qq, dm, kx, ix, jx, kx)
end subroutine bdfrdz

!=======================================================================
! the helper routines the engines are built from, with the reference argument
! lists (device counterparts; the engines above do not call them)
!=======================================================================
! cans3d/hdmlw/mlwhalf.f:2-3
subroutine mlwhalf(u, un, du, dt, fx, dxi, dxim, ix, fy, dyi, dyim, jx, fz, dzi, dzim, kx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx, kx
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
  double precision, target, intent(in) :: fz(*)
  double precision, target, intent(in) :: dzi(*)
  double precision, target, intent(in) :: dzim(*)
  integer :: n
  n = ix * jx * kx
  call check(c_mlwhalf3d(dmap(u, n), dmap(un, n), dmap(du, n), dt, dmap(fx, n), dmap(dxi, ix), dmap(dxim, ix), &
       int(ix, c_int), dmap(fy, n), dmap(dyi, jx), dmap(dyim, jx), int(jx, c_int), dmap(fz, n), dmap(dzi, kx), &
       dmap(dzim, kx), int(kx, c_int), c_null_ptr), 'mlwhalf')
  call dunmap(u, 0); call dunmap(un, 1); call dunmap(du, 1); call dunmap(fx, 0)
  call dunmap(dxi, 0); call dunmap(dxim, 0); call dunmap(fy, 0); call dunmap(dyi, 0)
  call dunmap(dyim, 0); call dunmap(fz, 0); call dunmap(dzi, 0); call dunmap(dzim, 0)
end subroutine mlwhalf

! cans3d/hdmlw/mlwfull.f:2-3
subroutine mlwfull(du, dt, fx, dxi, ux0, ux1, ix, fy, dyi, uy0, uy1, jx, fz, dzi, uz0, uz1, kx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx, kx
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
  double precision, target, intent(in) :: fz(*)
  double precision, target, intent(in) :: dzi(*)
  double precision, target, intent(in) :: uz0(*)
  double precision, target, intent(in) :: uz1(*)
  integer :: n
  n = ix * jx * kx
  call check(c_mlwfull3d(dmap(du, n), dt, dmap(fx, n), dmap(dxi, ix), dmap(ux0, ix), dmap(ux1, ix), int(ix, c_int), &
       dmap(fy, n), dmap(dyi, jx), dmap(uy0, jx), dmap(uy1, jx), int(jx, c_int), dmap(fz, n), dmap(dzi, kx), &
       dmap(uz0, kx), dmap(uz1, kx), int(kx, c_int), c_null_ptr), 'mlwfull')
  call dunmap(du, 1); call dunmap(fx, 0); call dunmap(dxi, 0); call dunmap(ux0, 0)
  call dunmap(ux1, 0); call dunmap(fy, 0); call dunmap(dyi, 0); call dunmap(uy0, 0)
  call dunmap(uy1, 0); call dunmap(fz, 0); call dunmap(dzi, 0); call dunmap(uz0, 0)
  call dunmap(uz1, 0)
end subroutine mlwfull

! cans3d/hdmlw/mlwartv.f:2-3
subroutine mlwartv(u, du, dt, qx, dxi, dxim, ix, qy, dyi, dyim, jx, qz, dzi, dzim, kx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx, kx
  double precision, target, intent(in) :: u(*)
  double precision, target :: du(*)
  double precision, intent(in) :: dt
  double precision, target, intent(in) :: qx(*)
  double precision, target, intent(in) :: dxi(*)
  double precision, target, intent(in) :: dxim(*)
  double precision, target, intent(in) :: qy(*)
  double precision, target, intent(in) :: dyi(*)
  double precision, target, intent(in) :: dyim(*)
  double precision, target, intent(in) :: qz(*)
  double precision, target, intent(in) :: dzi(*)
  double precision, target, intent(in) :: dzim(*)
  integer :: n
  n = ix * jx * kx
  call check(c_mlwartv3d(dmap(u, n), dmap(du, n), dt, dmap(qx, n), dmap(dxi, ix), dmap(dxim, ix), int(ix, c_int), &
       dmap(qy, n), dmap(dyi, jx), dmap(dyim, jx), int(jx, c_int), dmap(qz, n), dmap(dzi, kx), dmap(dzim, kx), &
       int(kx, c_int), c_null_ptr), 'mlwartv')
  call dunmap(u, 0); call dunmap(du, 1); call dunmap(qx, 0); call dunmap(dxi, 0)
  call dunmap(dxim, 0); call dunmap(qy, 0); call dunmap(dyi, 0); call dunmap(dyim, 0)
  call dunmap(qz, 0); call dunmap(dzi, 0); call dunmap(dzim, 0)
end subroutine mlwartv

! cans3d/hdmlw/mlwsrch.f:2
subroutine mlwsrch(un, du, dt, s, ix, jx, kx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx, kx
  double precision, target :: un(*)
  double precision, target :: du(*)
  double precision, intent(in) :: dt
  double precision, target, intent(in) :: s(*)
  integer :: n
  n = ix * jx * kx
  call check(c_mlwsrch3d(dmap(un, n), dmap(du, n), dt, dmap(s, n), int(ix, c_int), int(jx, c_int), int(kx, c_int), &
       c_null_ptr), 'mlwsrch')
  call dunmap(un, 1); call dunmap(du, 1); call dunmap(s, 0)
end subroutine mlwsrch

! cans3d/hdmlw/mlwsrcf.f:2
subroutine mlwsrcf(du, dt, s, ux0, ux1, ix, uy0, uy1, jx, uz0, uz1, kx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx, kx
  double precision, target :: du(*)
  double precision, intent(in) :: dt
  double precision, target, intent(in) :: s(*)
  double precision, target, intent(in) :: ux0(*)
  double precision, target, intent(in) :: ux1(*)
  double precision, target, intent(in) :: uy0(*)
  double precision, target, intent(in) :: uy1(*)
  double precision, target, intent(in) :: uz0(*)
  double precision, target, intent(in) :: uz1(*)
  integer :: n
  n = ix * jx * kx
  call check(c_mlwsrcf3d(dmap(du, n), dt, dmap(s, n), dmap(ux0, ix), dmap(ux1, ix), int(ix, c_int), dmap(uy0, jx), &
       dmap(uy1, jx), int(jx, c_int), dmap(uz0, kx), dmap(uz1, kx), int(kx, c_int), c_null_ptr), 'mlwsrcf')
  call dunmap(du, 1); call dunmap(s, 0); call dunmap(ux0, 0); call dunmap(ux1, 0)
  call dunmap(uy0, 0); call dunmap(uy1, 0); call dunmap(uz0, 0); call dunmap(uz1, 0)
end subroutine mlwsrcf

! cans3d/hdroe/roeflux_m.f:6-7
subroutine roeflux_m(fro, fee, frx, fry, frz, fby, fbz, gm, row, prw, vxw, vyw, vzw, bxw, byw, bzw, ix, jx, kx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx, kx
  double precision, target :: fro(*)
  double precision, target :: fee(*)
  double precision, target :: frx(*)
  double precision, target :: fry(*)
  double precision, target :: frz(*)
  double precision, target :: fby(*)
  double precision, target :: fbz(*)
  double precision, intent(in) :: gm
  double precision, target, intent(in) :: row(*)
  double precision, target, intent(in) :: prw(*)
  double precision, target, intent(in) :: vxw(*)
  double precision, target, intent(in) :: vyw(*)
  double precision, target, intent(in) :: vzw(*)
  double precision, target, intent(in) :: bxw(*)
  double precision, target, intent(in) :: byw(*)
  double precision, target, intent(in) :: bzw(*)
  integer :: n
  n = ix * jx * kx
  call check(c_roeflux_m3d(dmap(fro, n), dmap(fee, n), dmap(frx, n), dmap(fry, n), dmap(frz, n), dmap(fby, n), dmap(fbz, n), &
       gm, dmap(row, 2*n), dmap(prw, 2*n), dmap(vxw, 2*n), dmap(vyw, 2*n), dmap(vzw, 2*n), dmap(bxw, 2*n), &
       dmap(byw, 2*n), dmap(bzw, 2*n), int(ix, c_int), int(jx, c_int), int(kx, c_int), c_null_ptr), 'roeflux_m')
  call dunmap(fro, 1); call dunmap(fee, 1); call dunmap(frx, 1); call dunmap(fry, 1)
  call dunmap(frz, 1); call dunmap(fby, 1); call dunmap(fbz, 1); call dunmap(row, 0)
  call dunmap(prw, 0); call dunmap(vxw, 0); call dunmap(vyw, 0); call dunmap(vzw, 0)
  call dunmap(bxw, 0); call dunmap(byw, 0); call dunmap(bzw, 0)
end subroutine roeflux_m

! cans3d/hdroe/tvdminmod.f:4
subroutine tvdminmod(mdir, da, daw, ix, jx, kx)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: mdir, ix, jx, kx
  double precision, target, intent(in) :: da(*)
  double precision, target :: daw(*)
  integer :: n
  n = ix * jx * kx
  call check(c_tvdminmod3d(int(mdir, c_int), dmap(da, n), dmap(daw, 2*n), int(ix, c_int), int(jx, c_int), int(kx, c_int), &
       c_null_ptr), 'tvdminmod')
  call dunmap(da, 0); call dunmap(daw, 1)
end subroutine tvdminmod

! cansxx-nosupport/hdhlld/hlld_flux.f90:2-3
subroutine hlld_flux(row, prw, vxw, vyw, vzw, bx, byw, bzw, gm, ix, jx, kx, zfloor, fro, fee, frx, fry, frz, fby, fbz)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx, kx
  double precision, target, intent(in) :: row(*)
  double precision, target, intent(in) :: prw(*)
  double precision, target, intent(in) :: vxw(*)
  double precision, target, intent(in) :: vyw(*)
  double precision, target, intent(in) :: vzw(*)
  double precision, target, intent(in) :: bx(*)
  double precision, target, intent(in) :: byw(*)
  double precision, target, intent(in) :: bzw(*)
  double precision, intent(in) :: gm
  double precision, intent(in) :: zfloor
  double precision, target :: fro(*)
  double precision, target :: fee(*)
  double precision, target :: frx(*)
  double precision, target :: fry(*)
  double precision, target :: frz(*)
  double precision, target :: fby(*)
  double precision, target :: fbz(*)
  integer :: n
  n = ix * jx * kx
  call check(c_hlld_flux3d(dmap(row, 2*n), dmap(prw, 2*n), dmap(vxw, 2*n), dmap(vyw, 2*n), dmap(vzw, 2*n), dmap(bx, n), &
       dmap(byw, 2*n), dmap(bzw, 2*n), gm, int(ix, c_int), int(jx, c_int), int(kx, c_int), zfloor, &
       dmap(fro, n), dmap(fee, n), dmap(frx, n), dmap(fry, n), dmap(frz, n), dmap(fby, n), dmap(fbz, n), &
       c_null_ptr), 'hlld_flux')
  call dunmap(row, 0); call dunmap(prw, 0); call dunmap(vxw, 0); call dunmap(vyw, 0)
  call dunmap(vzw, 0); call dunmap(bx, 0); call dunmap(byw, 0); call dunmap(bzw, 0)
  call dunmap(fro, 1); call dunmap(fee, 1); call dunmap(frx, 1); call dunmap(fry, 1)
  call dunmap(frz, 1); call dunmap(fby, 1); call dunmap(fbz, 1)
end subroutine hlld_flux

! cansxx-nosupport/hdhlld/hlld_mh.f:2-4
subroutine hlld_mh(rof, prf, vxf, vyf, vzf, bxf, byf, bzf, ron, een, rxn, ryn, rzn, bxn, byn, bzn, gm, dt, dx, dy, dz, ix, jx, kx, mfdim, mstage)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: ix, jx, kx, mstage
  double precision, target, intent(in) :: rof(*)
  double precision, target, intent(in) :: prf(*)
  double precision, target, intent(in) :: vxf(*)
  double precision, target, intent(in) :: vyf(*)
  double precision, target, intent(in) :: vzf(*)
  double precision, target, intent(in) :: bxf(*)
  double precision, target, intent(in) :: byf(*)
  double precision, target, intent(in) :: bzf(*)
  double precision, target :: ron(*)
  double precision, target :: een(*)
  double precision, target :: rxn(*)
  double precision, target :: ryn(*)
  double precision, target :: rzn(*)
  double precision, target :: bxn(*)
  double precision, target :: byn(*)
  double precision, target :: bzn(*)
  double precision, intent(in) :: gm
  double precision, intent(in) :: dt
  double precision, target, intent(in) :: dx(*)
  double precision, target, intent(in) :: dy(*)
  double precision, target, intent(in) :: dz(*)
  integer, target, intent(in) :: mfdim(3)
  integer :: n
  n = ix * jx * kx
  call check(c_hlld_mh3d(dmap(rof, n), dmap(prf, n), dmap(vxf, n), dmap(vyf, n), dmap(vzf, n), dmap(bxf, n), dmap(byf, n), &
       dmap(bzf, n), dmap(ron, n), dmap(een, n), dmap(rxn, n), dmap(ryn, n), dmap(rzn, n), dmap(bxn, n), &
       dmap(byn, n), dmap(bzn, n), gm, dt, dmap(dx, ix), dmap(dy, jx), dmap(dz, kx), int(ix, c_int), &
       int(jx, c_int), int(kx, c_int), c_loc(mfdim(1)), int(mstage, c_int), c_null_ptr), 'hlld_mh')
  call dunmap(rof, 0); call dunmap(prf, 0); call dunmap(vxf, 0); call dunmap(vyf, 0)
  call dunmap(vzf, 0); call dunmap(bxf, 0); call dunmap(byf, 0); call dunmap(bzf, 0)
  call dunmap(ron, 1); call dunmap(een, 1); call dunmap(rxn, 1); call dunmap(ryn, 1)
  call dunmap(rzn, 1); call dunmap(bxn, 1); call dunmap(byn, 1); call dunmap(bzn, 1)
  call dunmap(dx, 0); call dunmap(dy, 0); call dunmap(dz, 0)
end subroutine hlld_mh

