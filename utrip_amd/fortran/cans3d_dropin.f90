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
