!=======================================================================
! cans_gpu_iface -- ISO_C_BINDING interface to libcans_gpu.so
! (include/cans_gpu.h) and the helpers the drop-in wrappers share.
!
! The wrappers keep the CANS calling convention (implicit interfaces,
! arrays by address, explicit shape with the ghost layers).  Each array is
! mirrored on the GPU through cans_gpu_map/cans_gpu_unmap: in the default
! mode every call copies its arrays in and the modified ones out; after
! call cans_gpu_resident(1) the arrays stay on the device between calls and
! call cans_gpu_sync() copies them back (before output).
!=======================================================================
module cans_gpu_iface
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: BD_FRE = 0, BD_SPP = 1, BD_SPN = 2, BD_SMP = 3, &
       BD_SMN = 4, BD_CNS = 5, BD_FRD = 6, BD_PER = 7

  interface
     function c_map(host, nbytes, copy_in) bind(C, name="cans_gpu_map")
       import :: c_ptr, c_size_t, c_int
       type(c_ptr), value :: host
       integer(c_size_t), value :: nbytes
       integer(c_int), value :: copy_in
       type(c_ptr) :: c_map
     end function c_map
     integer(c_int) function c_unmap(host, copy_out) bind(C, name="cans_gpu_unmap")
       import :: c_ptr, c_int
       type(c_ptr), value :: host
       integer(c_int), value :: copy_out
     end function c_unmap
     integer(c_int) function c_set_resident(on) bind(C, name="cans_gpu_set_resident")
       import :: c_int
       integer(c_int), value :: on
     end function c_set_resident
     integer(c_int) function c_sync_host() bind(C, name="cans_gpu_sync_host")
       import :: c_int
     end function c_sync_host
     integer(c_int) function c_invalidate(host) bind(C, name="cans_gpu_invalidate")
       import :: c_ptr, c_int
       type(c_ptr), value :: host
     end function c_invalidate
     integer(c_int) function c_forget(host) bind(C, name="cans_gpu_forget")
       import :: c_ptr, c_int
       type(c_ptr), value :: host
     end function c_forget
     function c_last_error() bind(C, name="cans_gpu_last_error")
       import :: c_ptr
       type(c_ptr) :: c_last_error
     end function c_last_error

     ! ---------------------------------------------------------- 1D
     integer(c_int) function c_mlw1d(ro, pr, vx, vy, by, bx, bxm, dt, qav, gm, dx, dxm, ix, s) &
          bind(C, name="cans1d_mlw_m")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: ro, pr, vx, vy, by, bx, bxm, dx, dxm, s
       real(c_double), value :: dt, qav, gm
       integer(c_int), value :: ix
     end function c_mlw1d
     integer(c_int) function c_mlw1d_bg(ro, pr, vx, vy, by, bx, bxm, dt, qav, gm, gx, gxm, sc, dsc, scm, dscm, &
          rr, rrm, drr, drrm, dx, dxm, ix, s) bind(C, name="cans1d_mlw_m_bg")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: ro, pr, vx, vy, by, bx, bxm, gx, gxm, sc, dsc, scm, dscm, rr, rrm, drr, drrm, dx, dxm, s
       real(c_double), value :: dt, qav, gm
       integer(c_int), value :: ix
     end function c_mlw1d_bg
     integer(c_int) function c_trq1d(vy, dt, trq, rr, ix, mseed, s) bind(C, name="cans1d_trqalfven")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: vy, trq, rr, s
       real(c_double), value :: dt
       integer(c_int), value :: ix
       integer(c_int) :: mseed
     end function c_trq1d
     integer(c_int) function c_chkdav1d(n_events, da, vx, floor, ix, s) bind(C, name="cans1d_chkdav")
       import :: c_ptr, c_double, c_int
       integer(c_int) :: n_events
       type(c_ptr), value :: da, vx, s
       real(c_double), value :: floor
       integer(c_int), value :: ix
     end function c_chkdav1d
     integer(c_int) function c_roe1d(ro, pr, vx, vy, by, bx, bxm, dt, gm, dx, ix, s) &
          bind(C, name="cans1d_roe_m")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: ro, pr, vx, vy, by, bx, bxm, dx, s
       real(c_double), value :: dt, gm
       integer(c_int), value :: ix
     end function c_roe1d
     integer(c_int) function c_cfl1d(dt, safety, dtmin, merr, gm, bx, ro, pr, vx, vy, by, dx, ix, s) &
          bind(C, name="cans1d_cfl_m")
       import :: c_ptr, c_double, c_int
       real(c_double) :: dt
       integer(c_int) :: merr
       real(c_double), value :: safety, dtmin, gm
       type(c_ptr), value :: bx, ro, pr, vx, vy, by, dx, s
       integer(c_int), value :: ix
     end function c_cfl1d
     integer(c_int) function c_bnd1d(kind, mbnd, margin, qq, q0, dxm, ix, s) bind(C, name="cans1d_bnd")
       import :: c_ptr, c_double, c_int
       integer(c_int), value :: kind, mbnd, margin, ix
       type(c_ptr), value :: qq, dxm, s
       real(c_double), value :: q0
     end function c_bnd1d

     integer(c_int) function c_mlw1d_m3(ro, pr, vx, vy, vz, by, bz, bx, bxm, dt, qav, gm, dx, dxm, ix, s) &
          bind(C, name="cans1d_mlw_m3")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: ro, pr, vx, vy, vz, by, bz, bx, bxm, dx, dxm, s
       real(c_double), value :: dt, qav, gm
       integer(c_int), value :: ix
     end function c_mlw1d_m3
     integer(c_int) function c_roe1d_m3(ro, pr, vx, vy, vz, by, bz, bx, bxm, dt, gm, dx, ix, s) &
          bind(C, name="cans1d_roe_m3")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: ro, pr, vx, vy, vz, by, bz, bx, bxm, dx, s
       real(c_double), value :: dt, gm
       integer(c_int), value :: ix
     end function c_roe1d_m3
     integer(c_int) function c_cfl1d_m3(dt, safety, dtmin, merr, gm, bx, ro, pr, vx, vy, by, vz, bz, dx, ix, s) &
          bind(C, name="cans1d_cfl_m3")
       import :: c_ptr, c_double, c_int
       real(c_double) :: dt
       integer(c_int) :: merr
       real(c_double), value :: safety, dtmin, gm
       type(c_ptr), value :: bx, ro, pr, vx, vy, by, vz, bz, dx, s
       integer(c_int), value :: ix
     end function c_cfl1d_m3

     ! ---------------------------------------------------------- 2D
     integer(c_int) function c_mlw2d(ro, pr, vx, vy, bx, by, az, dt, qav, gm, dx, dxm, ix, dy, dym, jx, s) &
          bind(C, name="cans2d_mlw_m")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: ro, pr, vx, vy, bx, by, az, dx, dxm, dy, dym, s
       real(c_double), value :: dt, qav, gm
       integer(c_int), value :: ix, jx
     end function c_mlw2d
     integer(c_int) function c_roe2d(ro, pr, vx, vy, bx, by, az, dt, gm, dx, ix, dy, jx, s) &
          bind(C, name="cans2d_roe_m")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: ro, pr, vx, vy, bx, by, az, dx, dy, s
       real(c_double), value :: dt, gm
       integer(c_int), value :: ix, jx
     end function c_roe2d
     integer(c_int) function c_mlw2d_m3(ro, pr, vx, vy, vz, bx, by, bz, az, dt, qav, gm, dx, dxm, ix, dy, dym, &
          jx, s) bind(C, name="cans2d_mlw_m3")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: ro, pr, vx, vy, vz, bx, by, bz, az, dx, dxm, dy, dym, s
       real(c_double), value :: dt, qav, gm
       integer(c_int), value :: ix, jx
     end function c_mlw2d_m3
     integer(c_int) function c_mlw2d_c(ro, pr, vx, vz, bx, bz, ay, dt, qav, gm, x, xm, dx, dxm, ix, dz, dzm, &
          jx, s) bind(C, name="cans2d_mlw_m_c")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: ro, pr, vx, vz, bx, bz, ay, x, xm, dx, dxm, dz, dzm, s
       real(c_double), value :: dt, qav, gm
       integer(c_int), value :: ix, jx
     end function c_mlw2d_c
     integer(c_int) function c_mlw2d_s(ro, pr, vx, vy, bx, by, az, dt, qav, gm, x, xm, dx, dxm, ix, y, ym, &
          dy, dym, jx, s) bind(C, name="cans2d_mlw_m_s")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: ro, pr, vx, vy, bx, by, az, x, xm, dx, dxm, y, ym, dy, dym, s
       real(c_double), value :: dt, qav, gm
       integer(c_int), value :: ix, jx
     end function c_mlw2d_s
     integer(c_int) function c_cfl2d_s(dt, safety, dtmin, merr, gm, ro, pr, vx, vy, bx, by, x, dx, ix, dy, jx, s) &
          bind(C, name="cans2d_cfl_m_s")
       import :: c_ptr, c_double, c_int
       real(c_double) :: dt
       integer(c_int) :: merr
       real(c_double), value :: safety, dtmin, gm
       type(c_ptr), value :: ro, pr, vx, vy, bx, by, x, dx, dy, s
       integer(c_int), value :: ix, jx
     end function c_cfl2d_s
     integer(c_int) function c_roe2d_m3(ro, pr, vx, vy, vz, bx, by, bz, az, dt, gm, dx, ix, dy, jx, s) &
          bind(C, name="cans2d_roe_m3")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: ro, pr, vx, vy, vz, bx, by, bz, az, dx, dy, s
       real(c_double), value :: dt, gm
       integer(c_int), value :: ix, jx
     end function c_roe2d_m3
     integer(c_int) function c_cfl2d_m3(dt, safety, dtmin, merr, gm, ro, pr, vx, vy, vz, bx, by, bz, dx, ix, dy, &
          jx, s) bind(C, name="cans2d_cfl_m3")
       import :: c_ptr, c_double, c_int
       real(c_double) :: dt
       integer(c_int) :: merr
       real(c_double), value :: safety, dtmin, gm
       type(c_ptr), value :: ro, pr, vx, vy, vz, bx, by, bz, dx, dy, s
       integer(c_int), value :: ix, jx
     end function c_cfl2d_m3
     integer(c_int) function c_cfl2d(dt, safety, dtmin, merr, gm, ro, pr, vx, vy, bx, by, dx, ix, dy, jx, s) &
          bind(C, name="cans2d_cfl_m")
       import :: c_ptr, c_double, c_int
       real(c_double) :: dt
       integer(c_int) :: merr
       real(c_double), value :: safety, dtmin, gm
       type(c_ptr), value :: ro, pr, vx, vy, bx, by, dx, dy, s
       integer(c_int), value :: ix, jx
     end function c_cfl2d
     integer(c_int) function c_bnd2d(dir, kind, mbnd, margin, qq, q0, dm, ix, jx, s) bind(C, name="cans2d_bnd")
       import :: c_ptr, c_double, c_int
       integer(c_int), value :: dir, kind, mbnd, margin, ix, jx
       type(c_ptr), value :: qq, dm, s
       real(c_double), value :: q0
     end function c_bnd2d

     ! ---------------------------------------------------------- 3D
     integer(c_int) function c_mlw3d(ro, pr, vx, vy, vz, bx, by, bz, dt, qav, gm, dx, dxm, ix, dy, dym, jx, &
          dz, dzm, kx, s) bind(C, name="cans3d_mlw_m")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: ro, pr, vx, vy, vz, bx, by, bz, dx, dxm, dy, dym, dz, dzm, s
       real(c_double), value :: dt, qav, gm
       integer(c_int), value :: ix, jx, kx
     end function c_mlw3d
     integer(c_int) function c_mlw3d_c(ro, pr, vx, vy, vz, bx, by, bz, dt, qav, gm, x, xm, dx, dxm, ix, &
          dy, dym, jx, dz, dzm, kx, s) bind(C, name="cans3d_mlw_m_c")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: ro, pr, vx, vy, vz, bx, by, bz, x, xm, dx, dxm, dy, dym, dz, dzm, s
       real(c_double), value :: dt, qav, gm
       integer(c_int), value :: ix, jx, kx
     end function c_mlw3d_c
     integer(c_int) function c_roe3d(ro, pr, vx, vy, vz, bx, by, bz, dt, gm, dx, ix, dy, jx, dz, kx, s) &
          bind(C, name="cans3d_roe_m")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: ro, pr, vx, vy, vz, bx, by, bz, dx, dy, dz, s
       real(c_double), value :: dt, gm
       integer(c_int), value :: ix, jx, kx
     end function c_roe3d
     integer(c_int) function c_cfl3d(dt, safety, dtmin, merr, gm, ro, pr, vx, vy, vz, bx, by, bz, &
          dx, ix, dy, jx, dz, kx, s) bind(C, name="cans3d_cfl_m")
       import :: c_ptr, c_double, c_int
       real(c_double) :: dt
       integer(c_int) :: merr
       real(c_double), value :: safety, dtmin, gm
       type(c_ptr), value :: ro, pr, vx, vy, vz, bx, by, bz, dx, dy, dz, s
       integer(c_int), value :: ix, jx, kx
     end function c_cfl3d
     integer(c_int) function c_cfl3d_s(dt, safety, dtmin, merr, gm, ro, pr, vx, vy, vz, bx, by, bz, &
          x, dx, ix, y, dy, jx, dz, kx, s) bind(C, name="cans3d_cfl_m_s")
       import :: c_ptr, c_double, c_int
       real(c_double) :: dt
       integer(c_int) :: merr
       real(c_double), value :: safety, dtmin, gm
       type(c_ptr), value :: ro, pr, vx, vy, vz, bx, by, bz, x, dx, y, dy, dz, s
       integer(c_int), value :: ix, jx, kx
     end function c_cfl3d_s
     integer(c_int) function c_mlw3d_s(ro, pr, vx, vy, vz, bx, by, bz, dt, qav, gm, x, xm, dx, dxm, ix, &
          y, ym, dy, dym, jx, dz, dzm, kx, s) bind(C, name="cans3d_mlw_m_s")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: ro, pr, vx, vy, vz, bx, by, bz, x, xm, dx, dxm, y, ym, dy, dym, dz, dzm, s
       real(c_double), value :: dt, qav, gm
       integer(c_int), value :: ix, jx, kx
     end function c_mlw3d_s
     integer(c_int) function c_bnd3d(dir, kind, mbnd, margin, qq, q0, dm, ix, jx, kx, s) bind(C, name="cans3d_bnd")
       import :: c_ptr, c_double, c_int
       integer(c_int), value :: dir, kind, mbnd, margin, ix, jx, kx
       type(c_ptr), value :: qq, dm, s
       real(c_double), value :: q0
     end function c_bnd3d
     integer(c_int) function c_hlld(ro, pr, vx, vy, vz, bx, by, bz, gm, mu, dt, dx, dy, dz, ix, jx, kx, mfdim, s) &
          bind(C, name="cans3d_hlld_engine")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: ro, pr, vx, vy, vz, bx, by, bz, dx, dy, dz, mfdim, s
       real(c_double), value :: gm, mu, dt
       integer(c_int), value :: ix, jx, kx
     end function c_hlld
     ! ---------------------------------------------------------- helper routines
     integer(c_int) function c_mlwhalf1d(u, un, du, dt, f, dxi, dxim, ix, strm) &
          bind(C, name="cans1d_mlwhalf")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: u, un, du, f, dxi, dxim, strm
       real(c_double), value :: dt
       integer(c_int), value :: ix
     end function c_mlwhalf1d
     integer(c_int) function c_mlwfull1d(du, dt, f, dxi, ix, strm) &
          bind(C, name="cans1d_mlwfull")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: du, f, dxi, strm
       real(c_double), value :: dt
       integer(c_int), value :: ix
     end function c_mlwfull1d
     integer(c_int) function c_mlwartv1d(u, du, dt, qx, dxi, dxim, ix, strm) &
          bind(C, name="cans1d_mlwartv")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: u, du, qx, dxi, dxim, strm
       real(c_double), value :: dt
       integer(c_int), value :: ix
     end function c_mlwartv1d
     integer(c_int) function c_mlwsrch1d(un, du, dt, s, ix, strm) &
          bind(C, name="cans1d_mlwsrch")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: un, du, s, strm
       real(c_double), value :: dt
       integer(c_int), value :: ix
     end function c_mlwsrch1d
     integer(c_int) function c_mlwsrcf1d(du, dt, s, ux0, ux1, ix, strm) &
          bind(C, name="cans1d_mlwsrcf")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: du, s, ux0, ux1, strm
       real(c_double), value :: dt
       integer(c_int), value :: ix
     end function c_mlwsrcf1d
     integer(c_int) function c_roeflux_m1d(fro, fee, frx, fry, fby, gm, row, prw, vxw, vyw, bxw, byw, ix, strm) &
          bind(C, name="cans1d_roeflux_m")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: fro, fee, frx, fry, fby, row, prw, vxw, vyw, bxw, byw, strm
       real(c_double), value :: gm
       integer(c_int), value :: ix
     end function c_roeflux_m1d
     integer(c_int) function c_tvdminmod1d(da, daw, ix, strm) &
          bind(C, name="cans1d_tvdminmod")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: da, daw, strm
       integer(c_int), value :: ix
     end function c_tvdminmod1d
     integer(c_int) function c_mlwhalf2d(u, un, du, dt, fx, dxi, dxim, ix, fy, dyi, dyim, jx, strm) &
          bind(C, name="cans2d_mlwhalf")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: u, un, du, fx, dxi, dxim, fy, dyi, dyim, strm
       real(c_double), value :: dt
       integer(c_int), value :: ix, jx
     end function c_mlwhalf2d
     integer(c_int) function c_mlwfull2d(du, dt, fx, dxi, ux0, ux1, ix, fy, dyi, uy0, uy1, jx, strm) &
          bind(C, name="cans2d_mlwfull")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: du, fx, dxi, ux0, ux1, fy, dyi, uy0, uy1, strm
       real(c_double), value :: dt
       integer(c_int), value :: ix, jx
     end function c_mlwfull2d
     integer(c_int) function c_mlwartv2d(u, du, dt, qx, dxi, dxim, ix, qy, dyi, dyim, jx, strm) &
          bind(C, name="cans2d_mlwartv")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: u, du, qx, dxi, dxim, qy, dyi, dyim, strm
       real(c_double), value :: dt
       integer(c_int), value :: ix, jx
     end function c_mlwartv2d
     integer(c_int) function c_mlwsrch2d(un, du, dt, s, ix, jx, strm) &
          bind(C, name="cans2d_mlwsrch")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: un, du, s, strm
       real(c_double), value :: dt
       integer(c_int), value :: ix, jx
     end function c_mlwsrch2d
     integer(c_int) function c_mlwsrcf2d(du, dt, s, ux0, ux1, ix, uy0, uy1, jx, strm) &
          bind(C, name="cans2d_mlwsrcf")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: du, s, ux0, ux1, uy0, uy1, strm
       real(c_double), value :: dt
       integer(c_int), value :: ix, jx
     end function c_mlwsrcf2d
     integer(c_int) function c_roeflux_m2d(fro, fee, frx, fry, fby, gm, row, prw, vxw, vyw, bxw, byw, ix, jx, strm) &
          bind(C, name="cans2d_roeflux_m")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: fro, fee, frx, fry, fby, row, prw, vxw, vyw, bxw, byw, strm
       real(c_double), value :: gm
       integer(c_int), value :: ix, jx
     end function c_roeflux_m2d
     integer(c_int) function c_tvdminmod2d(mdir, da, daw, ix, jx, strm) &
          bind(C, name="cans2d_tvdminmod")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: da, daw, strm
       integer(c_int), value :: mdir, ix, jx
     end function c_tvdminmod2d
     integer(c_int) function c_mlwhalf3d(u, un, du, dt, fx, dxi, dxim, ix, fy, dyi, dyim, jx, fz, dzi, dzim, kx, strm) &
          bind(C, name="cans3d_mlwhalf")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: u, un, du, fx, dxi, dxim, fy, dyi, dyim, fz, dzi, dzim, strm
       real(c_double), value :: dt
       integer(c_int), value :: ix, jx, kx
     end function c_mlwhalf3d
     integer(c_int) function c_mlwfull3d(du, dt, fx, dxi, ux0, ux1, ix, fy, dyi, uy0, uy1, jx, fz, dzi, uz0, uz1, kx, strm) &
          bind(C, name="cans3d_mlwfull")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: du, fx, dxi, ux0, ux1, fy, dyi, uy0, uy1, fz, dzi, uz0, uz1, strm
       real(c_double), value :: dt
       integer(c_int), value :: ix, jx, kx
     end function c_mlwfull3d
     integer(c_int) function c_mlwartv3d(u, du, dt, qx, dxi, dxim, ix, qy, dyi, dyim, jx, qz, dzi, dzim, kx, strm) &
          bind(C, name="cans3d_mlwartv")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: u, du, qx, dxi, dxim, qy, dyi, dyim, qz, dzi, dzim, strm
       real(c_double), value :: dt
       integer(c_int), value :: ix, jx, kx
     end function c_mlwartv3d
     integer(c_int) function c_mlwsrch3d(un, du, dt, s, ix, jx, kx, strm) &
          bind(C, name="cans3d_mlwsrch")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: un, du, s, strm
       real(c_double), value :: dt
       integer(c_int), value :: ix, jx, kx
     end function c_mlwsrch3d
     integer(c_int) function c_mlwsrcf3d(du, dt, s, ux0, ux1, ix, uy0, uy1, jx, uz0, uz1, kx, strm) &
          bind(C, name="cans3d_mlwsrcf")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: du, s, ux0, ux1, uy0, uy1, uz0, uz1, strm
       real(c_double), value :: dt
       integer(c_int), value :: ix, jx, kx
     end function c_mlwsrcf3d
     integer(c_int) function c_roeflux_m3d(fro, fee, frx, fry, frz, fby, fbz, gm, row, prw, vxw, vyw, vzw, bxw, byw, bzw, ix, jx, kx, strm) &
          bind(C, name="cans3d_roeflux_m")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: fro, fee, frx, fry, frz, fby, fbz, row, prw, vxw, vyw, vzw, bxw, byw, bzw, strm
       real(c_double), value :: gm
       integer(c_int), value :: ix, jx, kx
     end function c_roeflux_m3d
     integer(c_int) function c_tvdminmod3d(mdir, da, daw, ix, jx, kx, strm) &
          bind(C, name="cans3d_tvdminmod")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: da, daw, strm
       integer(c_int), value :: mdir, ix, jx, kx
     end function c_tvdminmod3d
     integer(c_int) function c_hlld_flux3d(row, prw, vxw, vyw, vzw, bx, byw, bzw, gm, ix, jx, kx, zfloor, fro, fee, frx, fry, frz, fby, fbz, strm) &
          bind(C, name="cans3d_hlld_flux")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: row, prw, vxw, vyw, vzw, bx, byw, bzw, fro, fee, frx, fry, frz, fby, fbz, strm
       real(c_double), value :: gm, zfloor
       integer(c_int), value :: ix, jx, kx
     end function c_hlld_flux3d
     integer(c_int) function c_hlld_mh3d(rof, prf, vxf, vyf, vzf, bxf, byf, bzf, ron, een, rxn, ryn, rzn, bxn, byn, bzn, gm, dt, dx, dy, dz, ix, jx, kx, mfdim, mstage, strm) &
          bind(C, name="cans3d_hlld_mh")
       import :: c_ptr, c_double, c_int
       type(c_ptr), value :: rof, prf, vxf, vyf, vzf, bxf, byf, bzf, ron, een, rxn, ryn, rzn, bxn, byn, bzn, dx, dy, dz, mfdim, strm
       real(c_double), value :: gm, dt
       integer(c_int), value :: ix, jx, kx, mstage
     end function c_hlld_mh3d
  end interface

contains

  ! device mirror of a host array of n doubles (copy in unless resident)
  function dmap(a, n) result(p)
    integer, intent(in) :: n
    double precision, target, intent(in) :: a(*)
    type(c_ptr) :: p
    p = c_map(c_loc(a(1)), int(n, c_size_t) * 8_c_size_t, 1_c_int)
    if (.not. c_associated(p)) call fail(10002, 'cans_gpu_map')
  end function dmap

  ! release a mirror; copy_out = 1 for arrays the routine modified
  subroutine dunmap(a, copy_out)
    double precision, target, intent(in) :: a(*)
    integer, intent(in) :: copy_out
    integer(c_int) :: rc
    rc = c_unmap(c_loc(a(1)), int(copy_out, c_int))
    if (rc /= 0) call fail(rc, 'cans_gpu_unmap')
  end subroutine dunmap

  subroutine check(rc, who)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: who
    if (rc /= 0) call fail(rc, who)
  end subroutine check

  subroutine fail(rc, who)
    integer, intent(in) :: rc
    character(*), intent(in) :: who
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    type(c_ptr) :: m
    m = c_last_error()
    n = 0
    if (c_associated(m)) then
       call c_f_pointer(m, msg, [512])
       do while (n < 512)
          if (msg(n + 1) == c_null_char) exit
          n = n + 1
       end do
       write(6, '(a,a,i6,a,512a)') trim(who), ': error ', rc, ': ', msg(1:n)
    else
       write(6, '(a,a,i6)') trim(who), ': error ', rc
    end if
    error stop 10000
  end subroutine fail


  subroutine bnd1(kind, mbnd, margin, qq, q0, ix)
    implicit none
    integer, intent(in) :: kind, mbnd, margin, ix
    double precision, target :: qq(ix)
    double precision, intent(in) :: q0
    call check(c_bnd1d(int(kind, c_int), int(mbnd, c_int), int(margin, c_int), dmap(qq, ix), q0, c_null_ptr, &
         int(ix, c_int), c_null_ptr), 'bd*x')
    call dunmap(qq, 1)
  end subroutine bnd1

  subroutine bnd2(dir, kind, mbnd, margin, qq, q0, ix, jx)
    implicit none
    integer, intent(in) :: dir, kind, mbnd, margin, ix, jx
    double precision, target :: qq(ix, jx)
    double precision, intent(in) :: q0
    call check(c_bnd2d(int(dir, c_int), int(kind, c_int), int(mbnd, c_int), int(margin, c_int), dmap(qq, ix * jx), &
         q0, c_null_ptr, int(ix, c_int), int(jx, c_int), c_null_ptr), 'bd*')
    call dunmap(qq, 1)
  end subroutine bnd2

  subroutine bnd2d_frd(dir, mbnd, margin, qq, dm, nd, ix, jx)
    implicit none
    integer, intent(in) :: dir, mbnd, margin, nd, ix, jx
    double precision, target :: qq(ix, jx)
    double precision, target, intent(in) :: dm(nd)
    call check(c_bnd2d(int(dir, c_int), BD_FRD, int(mbnd, c_int), int(margin, c_int), dmap(qq, ix * jx), 0.0d0, &
         dmap(dm, nd), int(ix, c_int), int(jx, c_int), c_null_ptr), 'bdfrd')
    call dunmap(qq, 1); call dunmap(dm, 0)
  end subroutine bnd2d_frd

  subroutine unmap8(ro, pr, vx, vy, vz, bx, by, bz)
    implicit none
    double precision, target :: ro(*), pr(*), vx(*), vy(*), vz(*), bx(*), by(*), bz(*)
    call dunmap(ro, 1); call dunmap(pr, 1); call dunmap(vx, 1); call dunmap(vy, 1)
    call dunmap(vz, 1); call dunmap(bx, 1); call dunmap(by, 1); call dunmap(bz, 1)
  end subroutine unmap8

  subroutine bnd3(dir, kind, mbnd, margin, qq, q0, ix, jx, kx)
    implicit none
    integer, intent(in) :: dir, kind, mbnd, margin, ix, jx, kx
    double precision, target :: qq(*)
    double precision, intent(in) :: q0
    call check(c_bnd3d(int(dir, c_int), int(kind, c_int), int(mbnd, c_int), int(margin, c_int), &
         dmap(qq, ix * jx * kx), q0, c_null_ptr, int(ix, c_int), int(jx, c_int), int(kx, c_int), c_null_ptr), 'bd*')
    call dunmap(qq, 1)
  end subroutine bnd3

  subroutine bnd3_frd(dir, mbnd, margin, qq, dm, nd, ix, jx, kx)
    implicit none
    integer, intent(in) :: dir, mbnd, margin, nd, ix, jx, kx
    double precision, target :: qq(*)
    double precision, target, intent(in) :: dm(nd)
    call check(c_bnd3d(int(dir, c_int), BD_FRD, int(mbnd, c_int), int(margin, c_int), dmap(qq, ix * jx * kx), &
         0.0d0, dmap(dm, nd), int(ix, c_int), int(jx, c_int), int(kx, c_int), c_null_ptr), 'bdfrd')
    call dunmap(qq, 1); call dunmap(dm, 0)
  end subroutine bnd3_frd

end module cans_gpu_iface

!-----------------------------------------------------------------------
! residency control for modified main programs
!-----------------------------------------------------------------------
subroutine cans_gpu_resident(on)
  use cans_gpu_iface
  implicit none
  integer, intent(in) :: on
  call check(c_set_resident(int(on, c_int)), 'cans_gpu_resident')
end subroutine cans_gpu_resident

subroutine cans_gpu_sync()
  use cans_gpu_iface
  implicit none
  call check(c_sync_host(), 'cans_gpu_sync')
end subroutine cans_gpu_sync

subroutine cans_gpu_modified(a)
  use cans_gpu_iface
  implicit none
  double precision, target, intent(in) :: a(*)
  call check(c_invalidate(c_loc(a(1))), 'cans_gpu_modified')
end subroutine cans_gpu_modified

subroutine cans_gpu_forget(a)
  use cans_gpu_iface
  implicit none
  double precision, target, intent(in) :: a(*)
  call check(c_forget(c_loc(a(1))), 'cans_gpu_forget')
end subroutine cans_gpu_forget
