/*
 * cans_gpu.h -- C-ABI of libcans_gpu.so, the MI355X (gfx950) implementation
 * of the CANS MHD time step.
 *
 * Every entry point takes plain pointers, sizes and scalars.  Arrays are FP64,
 * Fortran column-major with the ghost layers included, exactly as the CANS
 * routines receive them: element (i,j,k) (1-based) is at
 * (i-1) + ix*((j-1) + jx*(k-1)).  `stream` is a hipStream_t (NULL = the
 * default stream).  Return value: 0 on success, CANS_ERR_* otherwise, with a
 * message from cans_gpu_last_error().
 *
 * Citations are to the reference routine each entry point replaces
 * (paths relative to cans_beta_vary/).  The Fortran drop-in libraries
 * libcans{1,2,3}d_gpu.so bind these entry points with ISO_C_BINDING and
 * export the reference's own subroutine names (mlw_m, cfl_m, bdperx, ...).
 */
#ifndef CANS_GPU_H
#define CANS_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CANS_ERR_ARG 10001 /* bad argument (abort class, >= 10000 as merr) */
#define CANS_ERR_HIP 10002 /* HIP runtime error */

/* boundary kinds: the CANS bc/ routines */
#define CANS_BD_FRE 0  /* bdfre{x,y,z}: cans1d/bc/bdfrex.f:30-40  zero gradient        */
#define CANS_BD_SPP 1  /* bdspp{x,y,z}: cans1d/bc/bdsppx.f:32-42  mirror at face, +    */
#define CANS_BD_SPN 2  /* bdspn{x,y,z}: cans1d/bc/bdspnx.f        mirror at face, -    */
#define CANS_BD_SMP 3  /* bdsmp{x,y,z}: cans1d/bc/bdsmpx.f        mirror at centre, +  */
#define CANS_BD_SMN 4  /* bdsmn{x,y,z}: cans1d/bc/bdsmnx.f:33-45  mirror at centre, -, wall 0 */
#define CANS_BD_CNS 5  /* bdcns{x,y,z}: cans1d/bc/bdcnsx.f:31-41  constant q0          */
#define CANS_BD_FRD 6  /* bdfrd{x,y,z}: cans1d/bc/bdfrdx.f:31-43  linear extrapolation */
#define CANS_BD_PER 7  /* bdper{x,y,z}: cans1d/bc/bdperx.f:31-40  periodic, both sides */
#define CANS_BD_NONE (-1)

const char *cans_gpu_last_error(void);
int cans_gpu_init(int device);
int cans_gpu_sync(void *stream);
/* Tuning, no counterpart in the reference: the marching face kernels (HLLD
 * y/z sweeps, 3D Roe y/z fluxes) cut each column of faces into segments so
 * that one launch has at least `waves` wavefronts (default 32768; fewer
 * waves = longer segments).  Results do not depend on it. */
int cans_gpu_set_march_waves(int waves);
/* Tuning, no counterpart in the reference: the Cartesian 3D MLW step
 * (mlw_m 3D) runs as one kernel marching in z with its cell and corner
 * planes in LDS (path 1, the default) or as three kernels over HBM scratch
 * (path 0).  Both give the same bits. */
int cans_gpu_set_mlw3d_path(int path);

/* ---------------------------------------------------------------- 1D
 * cans1d/hdmlw/mlw_m.f:2 mlw_m(ro,pr,vx,vy,by,bx,bxm,dt,qav,gm,dx,dxm,ix)
 * (fuses mlwhalf.f, mlwfull.f, mlwartv.f); device arrays, updated in place
 * on i=2..ix-1. */
int cans1d_mlw_m(double *ro, double *pr, double *vx, double *vy, double *by,
                 const double *bx, const double *bxm, double dt, double qav, double gm,
                 const double *dx, const double *dxm, int ix, void *stream);
/* cans1d_mlw_m without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as cans1d_mlw_m; its workspace is a static device
 * array of 9*CANS_NOALLOC_MAX_IX doubles (4.5 MiB per GPU, held until the
 * process ends), so ix > CANS_NOALLOC_MAX_IX returns CANS_ERR_ARG.  All calls
 * on one GPU share that array: calls that overlap in time on one GPU (other
 * streams, other threads) race. */
#define CANS_NOALLOC_MAX_IX 65536
int cans1d_mlw_m_noalloc(double *ro, double *pr, double *vx, double *vy, double *by,
                         const double *bx, const double *bxm, double dt, double qav, double gm,
                         const double *dx, const double *dxm, int ix, void *stream);
/* cans1d/hdroe/roe_m.f:2 roe_m(ro,pr,vx,vy,by,bx,bxm,dt,gm,dx,ix)
 * (fuses roeflux_m.f, tvdminmod.f); updated in place on i=3..ix-2. */
int cans1d_roe_m(double *ro, double *pr, double *vx, double *vy, double *by,
                 const double *bx, const double *bxm, double dt, double gm,
                 const double *dx, int ix, void *stream);
/* cans1d/common/cfl_m.f:2 cfl_m(dt,safety,dtmin,merr,gm,bx,ro,pr,vx,vy,by,dx,ix);
 * dt and merr are host pointers (the call synchronises the stream). */
int cans1d_cfl_m(double *dt, double safety, double dtmin, int *merr, double gm,
                 const double *bx, const double *ro, const double *pr, const double *vx,
                 const double *vy, const double *by, const double *dx, int ix, void *stream);
/* cans1d/hdmlw/mlw_m3.f:2 mlw_m3(ro,pr,vx,vy,vz,by,bz,bx,bxm,dt,qav,gm,dx,dxm,ix):
 * 1D MHD with three velocity and field components; updated in place on
 * i=2..ix-1. */
int cans1d_mlw_m3(double *ro, double *pr, double *vx, double *vy, double *vz, double *by, double *bz,
                  const double *bx, const double *bxm, double dt, double qav, double gm, const double *dx,
                  const double *dxm, int ix, void *stream);
/* cans1d/hdroe/roe_m3.f:2 roe_m3(ro,pr,vx,vy,vz,by,bz,bx,bxm,dt,gm,dx,ix) (fuses
 * roeflux_m3.f, tvdminmod.f); updated in place on i=3..ix-2. */
int cans1d_roe_m3(double *ro, double *pr, double *vx, double *vy, double *vz, double *by, double *bz,
                  const double *bx, const double *bxm, double dt, double gm, const double *dx, int ix,
                  void *stream);
/* cans1d/common/cfl_m3.f:2 cfl_m3(dt,safety,dtmin,merr,gm,bx,ro,pr,vx,vy,by,vz,bz,dx,ix) */
int cans1d_cfl_m3(double *dt, double safety, double dtmin, int *merr, double gm, const double *bx,
                  const double *ro, const double *pr, const double *vx, const double *vy, const double *by,
                  const double *vz, const double *bz, const double *dx, int ix, void *stream);
/* cans1d/bc/bd*x.f: kind = CANS_BD_*; mbnd = side (0 low, 1 high), or for
 * CANS_BD_PER margin0 with margin = margin1; q0 for bdcnsx, dxm for bdfrdx. */
int cans1d_bnd(int kind, int mbnd, int margin, double *qq, double q0, const double *dxm,
               int ix, void *stream);

/* ---------------------------------------------------------------- 2D
 * cans2d/hdmlw/mlw_m.f:2-3 mlw_m(ro,pr,vx,vy,bx,by,az,dt,qav,gm,dx,dxm,ix,dy,dym,jx)
 * (fuses mlwhalf.f, mlwfull.f, mlwartv.f, mlwsrch.f, mlwsrcf.f); updated in
 * place on i=2..ix-1, j=2..jx-1. */
int cans2d_mlw_m(double *ro, double *pr, double *vx, double *vy, double *bx, double *by,
                 double *az, double dt, double qav, double gm, const double *dx,
                 const double *dxm, int ix, const double *dy, const double *dym, int jx,
                 void *stream);
/* cans2d/hdroe/roe_m.f:2 roe_m(ro,pr,vx,vy,bx,by,az,dt,gm,dx,ix,dy,jx);
 * updated in place on i=3..ix-2, j=3..jx-2. */
int cans2d_roe_m(double *ro, double *pr, double *vx, double *vy, double *bx, double *by,
                 double *az, double dt, double gm, const double *dx, int ix,
                 const double *dy, int jx, void *stream);
/* cans2d/common/cfl_m.f:2-3 cfl_m(dt,safety,dtmin,merr,gm,ro,pr,vx,vy,bx,by,dx,ix,dy,jx) */
int cans2d_cfl_m(double *dt, double safety, double dtmin, int *merr, double gm,
                 const double *ro, const double *pr, const double *vx, const double *vy,
                 const double *bx, const double *by, const double *dx, int ix,
                 const double *dy, int jx, void *stream);
/* cans2d/bc/bd*{x,y}.f; dir 0 = x, 1 = y */
int cans2d_bnd(int dir, int kind, int mbnd, int margin, double *qq, double q0,
               const double *dm, int ix, int jx, void *stream);

/* cans2d/hdmlw/mlw_m3.f:2-3 mlw_m3(ro,pr,vx,vy,vz,bx,by,bz,az,dt,qav,gm,dx,dxm,ix,dy,dym,jx):
 * 2D MHD with vz, bz (fuses mlwhalf.f, mlwfull.f, mlwartv.f, mlwsrch.f,
 * mlwsrcf.f); updated in place on i=2..ix-1, j=2..jx-1. */
int cans2d_mlw_m3(double *ro, double *pr, double *vx, double *vy, double *vz, double *bx, double *by, double *bz,
                  double *az, double dt, double qav, double gm, const double *dx, const double *dxm, int ix,
                  const double *dy, const double *dym, int jx, void *stream);
/* cans2d/hdroe/roe_m3.f:2 roe_m3(ro,pr,vx,vy,vz,bx,by,bz,az,dt,gm,dx,ix,dy,jx);
 * updated in place on i=3..ix-2, j=3..jx-2. */
int cans2d_roe_m3(double *ro, double *pr, double *vx, double *vy, double *vz, double *bx, double *by, double *bz,
                  double *az, double dt, double gm, const double *dx, int ix, const double *dy, int jx,
                  void *stream);
/* cans2d/hdmlw/mlw_m_c.f:2-3 mlw_m_c(ro,pr,vx,vz,bx,bz,ay,dt,qav,gm,x,xm,dx,dxm,ix,dz,dzm,jx):
 * axisymmetric cylindrical (r, z) = (x, z) MLW with the potential Ay; x, xm
 * are the cell and face radii; updated in place on i=2..ix-1, j=2..jx-1. */
int cans2d_mlw_m_c(double *ro, double *pr, double *vx, double *vz, double *bx, double *bz, double *ay, double dt,
                   double qav, double gm, const double *x, const double *xm, const double *dx, const double *dxm,
                   int ix, const double *dz, const double *dzm, int jx, void *stream);
/* cans2d/hdmlw/mlw_m_s.f:1-2 mlw_m_s(ro,pr,vx,vy,bx,by,az,dt,qav,gm,x,xm,dx,dxm,ix,y,ym,dy,dym,jx):
 * spherical (r, theta) = (x, y) MLW with Az; x, xm cell and face radii, y, ym cell and face colatitudes
 * (sin, cos taken on the host with the C library); updated in place on i=2..ix-1, j=2..jx-1. */
int cans2d_mlw_m_s(double *ro, double *pr, double *vx, double *vy, double *bx, double *by, double *az, double dt,
                   double qav, double gm, const double *x, const double *xm, const double *dx, const double *dxm,
                   int ix, const double *y, const double *ym, const double *dy, const double *dym, int jx,
                   void *stream);
/* cans2d/common/cfl_m_s.f:1-2 cfl_m_s(dt,safety,dtmin,merr,gm,ro,pr,vx,vy,bx,by,x,dx,ix,dy,jx):
 * cfl_m with the widths dx, x dy. */
int cans2d_cfl_m_s(double *dt, double safety, double dtmin, int *merr, double gm, const double *ro,
                   const double *pr, const double *vx, const double *vy, const double *bx, const double *by,
                   const double *x, const double *dx, int ix, const double *dy, int jx, void *stream);
/* cans2d/common/cfl_m3.f:2-3 cfl_m3(dt,safety,dtmin,merr,gm,ro,pr,vx,vy,vz,bx,by,bz,dx,ix,dy,jx) */
int cans2d_cfl_m3(double *dt, double safety, double dtmin, int *merr, double gm, const double *ro,
                  const double *pr, const double *vx, const double *vy, const double *vz, const double *bx,
                  const double *by, const double *bz, const double *dx, int ix, const double *dy, int jx,
                  void *stream);

/* ---------------------------------------------------------------- 3D
 * cans3d/hdmlw/mlw_m.f:2-3
 * mlw_m(ro,pr,vx,vy,vz,bx,by,bz,dt,qav,gm,dx,dxm,ix,dy,dym,jx,dz,dzm,kx) */
int cans3d_mlw_m(double *ro, double *pr, double *vx, double *vy, double *vz, double *bx,
                 double *by, double *bz, double dt, double qav, double gm,
                 const double *dx, const double *dxm, int ix, const double *dy,
                 const double *dym, int jx, const double *dz, const double *dzm, int kx,
                 void *stream);
/* cans3d/hdmlw/mlw_m_c.f:2-3 mlw_m_c(ro,pr,vx,vy,vz,bx,by,bz,dt,qav,gm,x,xm,dx,dxm,ix,dy,dym,jx,
 * dz,dzm,kx): cylindrical (r, phi, z) = (x, y, z) MLW (md_mhdsn/hdmlw/main.f:179); x, xm are the cell
 * and face radii; updated in place on the interior. */
int cans3d_mlw_m_c(double *ro, double *pr, double *vx, double *vy, double *vz, double *bx, double *by,
                   double *bz, double dt, double qav, double gm, const double *x, const double *xm,
                   const double *dx, const double *dxm, int ix, const double *dy, const double *dym, int jx,
                   const double *dz, const double *dzm, int kx, void *stream);
/* cans3d/hdmlw/mlw_m_s.f:2-3 mlw_m_s(ro,pr,vx,vy,vz,bx,by,bz,dt,qav,gm,x,xm,dx,dxm,ix,y,ym,dy,dym,jx,
 * dz,dzm,kx): spherical (r, theta, phi) = (x, y, z) MLW; x, xm the cell and face radii, y, ym the cell and
 * face colatitudes (their sin, cos, tan are taken on the host with the C library); updated in place on the
 * interior. */
int cans3d_mlw_m_s(double *ro, double *pr, double *vx, double *vy, double *vz, double *bx, double *by,
                   double *bz, double dt, double qav, double gm, const double *x, const double *xm,
                   const double *dx, const double *dxm, int ix, const double *y, const double *ym,
                   const double *dy, const double *dym, int jx, const double *dz, const double *dzm, int kx,
                   void *stream);
/* cans3d/hdroe/roe_m.f:2 roe_m(ro,pr,vx,vy,vz,bx,by,bz,dt,gm,dx,ix,dy,jx,dz,kx) */
int cans3d_roe_m(double *ro, double *pr, double *vx, double *vy, double *vz, double *bx,
                 double *by, double *bz, double dt, double gm, const double *dx, int ix,
                 const double *dy, int jx, const double *dz, int kx, void *stream);
/* cans3d/common/cfl_m.f:2-3
 * cfl_m(dt,safety,dtmin,merr,gm,ro,pr,vx,vy,vz,bx,by,bz,dx,ix,dy,jx,dz,kx) */
int cans3d_cfl_m(double *dt, double safety, double dtmin, int *merr, double gm,
                 const double *ro, const double *pr, const double *vx, const double *vy,
                 const double *vz, const double *bx, const double *by, const double *bz,
                 const double *dx, int ix, const double *dy, int jx, const double *dz, int kx,
                 void *stream);
/* cans3d/common/cfl_m_s.f:1-2 cfl_m_s(dt,safety,dtmin,merr,gm,ro,pr,vx,vy,vz,bx,by,bz,x,dx,ix,y,dy,jx,dz,kx):
 * cfl_m with the widths dx, x dy, x sin(y) dz. */
int cans3d_cfl_m_s(double *dt, double safety, double dtmin, int *merr, double gm,
                   const double *ro, const double *pr, const double *vx, const double *vy,
                   const double *vz, const double *bx, const double *by, const double *bz,
                   const double *x, const double *dx, int ix, const double *y, const double *dy, int jx,
                   const double *dz, int kx, void *stream);
/* cans3d/bc/bd*{x,y,z}.f; dir 0 = x, 1 = y, 2 = z */
int cans3d_bnd(int dir, int kind, int mbnd, int margin, double *qq, double q0,
               const double *dm, int ix, int jx, int kx, void *stream);
/* HLLD MUSCL-Hancock engine, cansxx-nosupport/md_mhdshktb/hdhlld-x/engine.f:2
 * engine(ro,pr,vx,vy,vz,bx,by,bz,gm,mu,dt,dx,dxm,dy,dym,dz,dzm,ix,jx,kx,mfdim)
 * (fuses hlld_mh.f, hlld_flux.f90, inp_const.f, inp_tvdminmod.f, adddfdx.f);
 * dxm/dym/dzm are unused by the engine and not taken. */
int cans3d_hlld_engine(double *ro, double *pr, double *vx, double *vy, double *vz,
                       double *bx, double *by, double *bz, double gm, double mu, double dt,
                       const double *dx, const double *dy, const double *dz, int ix, int jx,
                       int kx, const int *mfdim, void *stream);

/* Mixed precision: the same engines with the face-flux arithmetic in FP32
 * (roeflux / hlld_flux evaluated in float) and everything else -- the
 * conserved-variable accumulation, the MUSCL reconstruction, the pressure
 * recovery and cfl_m -- in FP64.  No counterpart in the reference; for
 * callers that trade ~1e-6 relative error for FP32 VALU throughput. */
int cans3d_roe_m_mixed(double *ro, double *pr, double *vx, double *vy, double *vz, double *bx,
                       double *by, double *bz, double dt, double gm, const double *dx, int ix,
                       const double *dy, int jx, const double *dz, int kx, void *stream);
int cans3d_hlld_engine_mixed(double *ro, double *pr, double *vx, double *vy, double *vz,
                             double *bx, double *by, double *bz, double gm, double mu, double dt,
                             const double *dx, const double *dy, const double *dz, int ix, int jx,
                             int kx, const int *mfdim, void *stream);

/* ---------------------------------------------------------------- helper routines
 * The routines the engines are built from, with their reference argument
 * lists (device arrays, in place as the reference updates them).  The fused
 * engines above do not call them; they are there for callers that compose
 * their own engine from them.  w arrays of roeflux_m, hlld_flux and the daw
 * output of tvdminmod have a trailing dimension 2 (left state first). */
/* cans1d/hdmlw/mlwhalf.f:2 mlwhalf(u,un,du,dt,f,dxi,dxim,ix) */
int cans1d_mlwhalf(const double *u, double *un, double *du, double dt, const double *f, const double *dxi,
                   const double *dxim, int ix, void *stream);
/* cans1d/hdmlw/mlwfull.f:2 mlwfull(du,dt,f,dxi,ix) */
int cans1d_mlwfull(double *du, double dt, const double *f, const double *dxi, int ix, void *stream);
/* cans1d/hdmlw/mlwartv.f:2 mlwartv(u,du,dt,qx,dxi,dxim,ix) */
int cans1d_mlwartv(const double *u, double *du, double dt, const double *qx, const double *dxi,
                   const double *dxim, int ix, void *stream);
/* cans1d/hdmlw/mlwsrch.f:2 mlwsrch(un,du,dt,s,ix) */
int cans1d_mlwsrch(double *un, double *du, double dt, const double *s, int ix, void *stream);
/* cans1d/hdmlw/mlwsrcf.f:2 mlwsrcf(du,dt,s,ux0,ux1,ix) */
int cans1d_mlwsrcf(double *du, double dt, const double *s, const double *ux0, const double *ux1, int ix,
                   void *stream);
/* cans1d/hdroe/roeflux_m.f:2-3 roeflux_m(fro,fee,frx,fry,fby,gm,row,prw,vxw,vyw,bxw,byw,ix), faces 1..ix-1 */
int cans1d_roeflux_m(double *fro, double *fee, double *frx, double *fry, double *fby, double gm,
                     const double *row, const double *prw, const double *vxw, const double *vyw,
                     const double *bxw, const double *byw, int ix, void *stream);
/* cans1d/hdroe/tvdminmod.f:2 tvdminmod(da,daw,ix) */
int cans1d_tvdminmod(const double *da, double *daw, int ix, void *stream);
/* cans2d/hdmlw/mlwhalf.f:2 mlwhalf(u,un,du,dt,fx,dxi,dxim,ix,fy,dyi,dyim,jx) */
int cans2d_mlwhalf(const double *u, double *un, double *du, double dt, const double *fx, const double *dxi,
                   const double *dxim, int ix, const double *fy, const double *dyi, const double *dyim, int jx,
                   void *stream);
/* cans2d/hdmlw/mlwfull.f:2 mlwfull(du,dt,fx,dxi,ux0,ux1,ix,fy,dyi,uy0,uy1,jx) */
int cans2d_mlwfull(double *du, double dt, const double *fx, const double *dxi, const double *ux0,
                   const double *ux1, int ix, const double *fy, const double *dyi, const double *uy0,
                   const double *uy1, int jx, void *stream);
/* cans2d/hdmlw/mlwartv.f:2 mlwartv(u,du,dt,qx,dxi,dxim,ix,qy,dyi,dyim,jx) */
int cans2d_mlwartv(const double *u, double *du, double dt, const double *qx, const double *dxi,
                   const double *dxim, int ix, const double *qy, const double *dyi, const double *dyim, int jx,
                   void *stream);
/* cans2d/hdmlw/mlwsrch.f:2 mlwsrch(un,du,dt,s,ix,jx) */
int cans2d_mlwsrch(double *un, double *du, double dt, const double *s, int ix, int jx, void *stream);
/* cans2d/hdmlw/mlwsrcf.f:2 mlwsrcf(du,dt,s,ux0,ux1,ix,uy0,uy1,jx) */
int cans2d_mlwsrcf(double *du, double dt, const double *s, const double *ux0, const double *ux1, int ix,
                   const double *uy0, const double *uy1, int jx, void *stream);
/* cans2d/hdroe/roeflux_m.f:1-2 roeflux_m(fro,fee,frx,fry,fby,gm,row,prw,vxw,vyw,bxw,byw,ix,jx) */
int cans2d_roeflux_m(double *fro, double *fee, double *frx, double *fry, double *fby, double gm,
                     const double *row, const double *prw, const double *vxw, const double *vyw,
                     const double *bxw, const double *byw, int ix, int jx, void *stream);
/* cans2d/hdroe/tvdminmod.f:1 tvdminmod(mdir,da,daw,ix,jx) */
int cans2d_tvdminmod(int mdir, const double *da, double *daw, int ix, int jx, void *stream);
/* cans3d/hdmlw/mlwhalf.f:2-3 mlwhalf(u,un,du,dt,fx,dxi,dxim,ix,fy,dyi,dyim,jx,fz,dzi,dzim,kx) */
int cans3d_mlwhalf(const double *u, double *un, double *du, double dt, const double *fx, const double *dxi,
                   const double *dxim, int ix, const double *fy, const double *dyi, const double *dyim, int jx,
                   const double *fz, const double *dzi, const double *dzim, int kx, void *stream);
/* cans3d/hdmlw/mlwfull.f:2-3 mlwfull(du,dt,fx,dxi,ux0,ux1,ix,fy,dyi,uy0,uy1,jx,fz,dzi,uz0,uz1,kx) */
int cans3d_mlwfull(double *du, double dt, const double *fx, const double *dxi, const double *ux0,
                   const double *ux1, int ix, const double *fy, const double *dyi, const double *uy0,
                   const double *uy1, int jx, const double *fz, const double *dzi, const double *uz0,
                   const double *uz1, int kx, void *stream);
/* cans3d/hdmlw/mlwartv.f:2-3 mlwartv(u,du,dt,qx,dxi,dxim,ix,qy,dyi,dyim,jx,qz,dzi,dzim,kx) */
int cans3d_mlwartv(const double *u, double *du, double dt, const double *qx, const double *dxi,
                   const double *dxim, int ix, const double *qy, const double *dyi, const double *dyim, int jx,
                   const double *qz, const double *dzi, const double *dzim, int kx, void *stream);
/* cans3d/hdmlw/mlwsrch.f:2 mlwsrch(un,du,dt,s,ix,jx,kx) */
int cans3d_mlwsrch(double *un, double *du, double dt, const double *s, int ix, int jx, int kx, void *stream);
/* cans3d/hdmlw/mlwsrcf.f:2 mlwsrcf(du,dt,s,ux0,ux1,ix,uy0,uy1,jx,uz0,uz1,kx) */
int cans3d_mlwsrcf(double *du, double dt, const double *s, const double *ux0, const double *ux1, int ix,
                   const double *uy0, const double *uy1, int jx, const double *uz0, const double *uz1, int kx,
                   void *stream);
/* cans3d/hdroe/roeflux_m.f:6-7 roeflux_m(fro,fee,frx,fry,frz,fby,fbz,gm,row,prw,vxw,vyw,vzw,bxw,byw,bzw,ix,jx,kx) */
int cans3d_roeflux_m(double *fro, double *fee, double *frx, double *fry, double *frz, double *fby, double *fbz,
                     double gm, const double *row, const double *prw, const double *vxw, const double *vyw,
                     const double *vzw, const double *bxw, const double *byw, const double *bzw, int ix, int jx,
                     int kx, void *stream);
/* cans3d/hdroe/tvdminmod.f:4 tvdminmod(mdir,da,daw,ix,jx,kx) */
int cans3d_tvdminmod(int mdir, const double *da, double *daw, int ix, int jx, int kx, void *stream);
/* cansxx-nosupport/hdhlld/hlld_flux.f90:2-3
 * hlld_flux(row,prw,vxw,vyw,vzw,bx,byw,bzw,gm,ix,jx,kx,zfloor,fro,fee,frx,fry,frz,fby,fbz), every face */
int cans3d_hlld_flux(const double *row, const double *prw, const double *vxw, const double *vyw,
                     const double *vzw, const double *bx, const double *byw, const double *bzw, double gm, int ix,
                     int jx, int kx, double zfloor, double *fro, double *fee, double *frx, double *fry,
                     double *frz, double *fby, double *fbz, void *stream);
/* cansxx-nosupport/hdhlld/hlld_mh.f:2-4 hlld_mh(rof,prf,vxf,vyf,vzf,bxf,byf,bzf,ron,een,rxn,ryn,rzn,
 * bxn,byn,bzn,gm,dt,dx,dy,dz,ix,jx,kx,mfdim,mstage); mfdim is a host array of 3.  Where the reference
 * reads its unset automatic arrays (the last face along a direction, which reaches ghost cells only)
 * the result here is NaN. */
int cans3d_hlld_mh(const double *rof, const double *prf, const double *vxf, const double *vyf,
                   const double *vzf, const double *bxf, const double *byf, const double *bzf, double *ron,
                   double *een, double *rxn, double *ryn, double *rzn, double *bxn, double *byn, double *bzn,
                   double gm, double dt, const double *dx, const double *dy, const double *dz, int ix, int jx,
                   int kx, const int *mfdim, int mstage, void *stream);

/* ---------------------------------------------------------------- resident loop
 * A device clock holds t, ns, merr and dt of the loop in
 * cans2d/md_mhdshktb/hdmlw/main.f:150-206 so that steps run back to back with
 * no host round trip; read it back with cans_clock_read().  Step n of a run
 * reads its state from buffer set a (n even) or b (n odd) and writes the
 * other; after N issued steps the current state is in a if nexec is even.
 * No counterpart in the reference (which keeps these scalars on the host). */
int cans_clock_create(void **clk);
int cans_clock_destroy(void *clk);
int cans_clock_reset(void *clk, double t0, double tend, int nstop, double dtmin, void *stream);
int cans_clock_read(const void *clk, int nsteps_issued, double *t, int *ns, int *nexec,
                    int *merr, double *dt_last, void *stream);
/* nsteps of {cfl_m; t=t+dt; mlw_m; bnd} on 2D state (7 device pointers per
 * set: ro,pr,vx,vy,bx,by,az).  bc == NULL: periodic box of width `margin`
 * (ghost fill and next-step cfl fused into the engine kernel); otherwise
 * bc[(dir*2+side)*7+field] gives a CANS_BD_* kind per face and field.
 * metric_ws: device workspace of cans2d_metric_ws_size(ix,jx) doubles, filled
 * by the call with step0 == 0 and read by the later calls of the same run. */
int cans2d_mlw_run(double *const *a7, double *const *b7, void *clk, int step0, int nsteps,
                   double qav, double gm, double safety, const double *dx, const double *dxm,
                   int ix, const double *dy, const double *dym, int jx, const int *bc,
                   int margin, double *metric_ws, void *const *events, void *stream);
int cans2d_metric_ws_size(int ix, int jx);
/* Fill a metric workspace (the layout cans2d_mlw_run uses) from dx,dxm,dy,dym. */
int cans2d_metric_fill(const double *dx, const double *dxm, int ix, const double *dy, const double *dym, int jx,
                       double *metric_ws, void *stream);

/* Reverse mode of one step of the fused periodic 2D MLW loop
 * P^{n+1} = bnd(mlw_m(P^n, dt_n)), dt_n = cfl_m(P^n)  (cans2d/hdmlw/mlw_m.f:2-324,
 * cans2d/bc/bdper{x,y}.f, cans2d/common/cfl_m.f).  pn6 / pn1_6: ro,pr,vx,vy,bx,by
 * of P^n and P^{n+1} (device, all cells); lam_in6: dL/dP^{n+1} on all cells,
 * unchanged (the adjoint of the periodic fill is applied on the fly);
 * lam_out6 receives dL/dP^n on all cells (call cans2d_periodic_fold on the
 * last one).  The dependence of dt on P^n is included (split evenly over
 * cells whose cfl term ties at the minimum).  az carries no adjoint (no
 * other field reads it).  metric_ws from cans2d_metric_fill; ws of
 * cans2d_adj_ws_bytes.  No counterpart in the reference (which has no
 * adjoint). */
size_t cans2d_adj_ws_bytes(int ix, int jx);
int cans2d_mlw_adjoint_step(const double *const *pn6, const double *const *pn1_6, double *const *lam_in6,
                            double *const *lam_out6, double qav, double gm, double safety, const double *dx,
                            const double *dxm, int ix, const double *dy, const double *dym, int jx, int margin,
                            const double *metric_ws, void *ws, void *stream);
/* nsteps of the fused periodic 2D MLW loop from states[0..6] (ro..az), writing
 * P^{k} to states[7k .. 7k+6] (a host array of 7*(nsteps+1) device pointers),
 * on a clock reset by the caller (steps 0..nsteps-1); dt_log[k] (device)
 * receives the dt of step k.  For the adjoint's segment recompute. */
int cans2d_mlw_record(double *const *states, int nsteps, void *clk, double qav, double gm, double safety,
                      const double *dx, const double *dxm, int ix, const double *dy, const double *dym, int jx,
                      int margin, double *metric_ws, double *dt_log, void *stream);
/* cans2d_mlw_adjoint_step for n = nsteps-1 .. 0 over states and dt_log
 * recorded by cans2d_mlw_record: lam6 holds dL/dP^{nsteps} on entry and dL/dP^0 (all
 * cells, not folded) on return; tmp6 is scratch of the same shape. */
int cans2d_mlw_adjoint_run(const double *const *states, const double *dt_log, int nsteps, double *const *lam6,
                           double *const *tmp6, double qav, double gm, double safety, const double *dx,
                           const double *dxm, int ix, const double *dy, const double *dym, int jx, int margin,
                           const double *metric_ws, void *ws, void *stream);
/* adjoint of bdperx then bdpery on 6 fields: ghost lam added to its source, cleared */
/* Reverse mode of the periodic 2D Roe-MUSCL loop (cans2d/hdroe/roe_m.f:2-208,
 * roeflux_m.f, tvdminmod.f, bdperx/bdpery, cfl_m; ghost width margin >= 3),
 * no counterpart in the reference: states holds 7*(nsteps+1) device
 * pointers (ro,pr,vx,vy,bx,by,az of every state of the trajectory, as the
 * resident Roe run or cans2d_roe_record wrote them), half NULL or the 6*nsteps
 * recorded half-step fields (cans2d_roe_record; else the step's first half is
 * recomputed), dt_log the nsteps dt (device); lam6 = dL/d
 * (ro,pr,vx,vy,bx,by) of the final state on entry, of the initial state on
 * return (not yet folded onto the interior: call cans2d_periodic_fold);
 * tmp6 a second set of the same size.  The face-flux derivative is the
 * generated reverse sweep of the kernels' roeflux12.  ws:
 * cans2d_roe_adj_ws_bytes. */
size_t cans2d_roe_adj_ws_bytes(int ix, int jx);
int cans2d_roe_adjoint_run(const double *const *states, const double *const *half, const double *dt_log, int nsteps,
                           double *const *lam6, double *const *tmp6, double gm, double safety, const double *dx,
                           const double *dy, int ix, int jx, int margin, void *ws, void *stream);
/* The periodic 2D Roe-MUSCL loop recorded for cans2d_roe_adjoint_run: nsteps
 * of cfl_m, the clock (clk: cans_clock_create, reset by the caller), roe_m and
 * bdperx/bdpery, state n+1 written to states[7(n+1) ..] (states[0..6] is the
 * initial state), the step's half-step primitives (ro pr vx vy bx by) to
 * half[6n ..] and its dt to dt_log[n] (device).  ws: cans_run_ws_bytes(2,
 * CANS_ENGINE_ROE, ix, jx, 1).  No counterpart in the reference. */
int cans2d_roe_record(double *const *states, double *const *half, int nsteps, void *clk, double gm, double safety,
                      const double *dx, const double *dy, int ix, int jx, int margin, void *ws, double *dt_log,
                      void *stream);
int cans2d_periodic_fold(double *const *lam6, int ix, int jx, int margin, void *stream);

/* Engines of the resident run */
#define CANS_ENGINE_MLW 0   /* cans{1,2,3}d/hdmlw/mlw_m.f */
#define CANS_ENGINE_ROE 1   /* cans{1,2,3}d/hdroe/roe_m.f */
#define CANS_ENGINE_HLLD 2  /* cansxx-nosupport/md_mhdshktb/hdhlld-x/engine.f (3D only) */
#define CANS_ENGINE_MLW3 3  /* cans{1,2}d/hdmlw/mlw_m3.f with cfl_m3 (1D/2D only) */
#define CANS_ENGINE_ROE3 4  /* cans{1,2}d/hdroe/roe_m3.f with cfl_m3 (1D/2D only) */
#define CANS_ENGINE_MLWC 5  /* cans{2,3}d/hdmlw/mlw_m_c.f, cylindrical (2D (r,z), 3D (r,phi,z); x, xm in bx1d, bxm1d) */
#define CANS_ENGINE_MLWS 6  /* cans{2,3}d/hdmlw/mlw_m_s.f + common/cfl_m_s.f, spherical (2D (r,theta) with az,
                             * 3D (r,theta,phi); bx1d = the device array x(ix), xm(ix), y(jx), ym(jx) back to
                             * back; bxm1d unused) */

/* Device workspace (bytes) a cans_run of this shape needs. */
size_t cans_run_ws_bytes(int dim, int engine, int ix, int jx, int kx);
/* nsteps of the CANS loop {cfl_m; t=t+dt; engine; bnd} for any dimension and
 * engine (mixed != 0: FP32 face fluxes for Roe/HLLD).  Field sets a/b:
 * 1D ro,pr,vx,vy,by (bx1d, bxm1d separate); 2D ro,pr,vx,vy,bx,by,az;
 * 3D ro,pr,vx,vy,vz,bx,by,bz.  bc[(dir*2+side)*nf+field] = CANS_BD_* kind
 * (NULL only for the fused periodic 2D MLW path).  mu is used by HLLD.
 * ws (cans_run_ws_bytes) holds the grid metric, filled by the call with
 * step0 == 0 and read by the later calls of the same run. */
int cans_run(int dim, int engine, int mixed, double *const *a, double *const *b, const double *bx1d,
             const double *bxm1d, void *clk, int step0, int nsteps, double qav, double gm, double mu,
             double safety, const double *dx, const double *dxm, int ix, const double *dy,
             const double *dym, int jx, const double *dz, const double *dzm, int kx, const int *bc,
             int margin, void *ws, void *const *events, void *stream);

/* ---------------------------------------------------------------- batched 1D tubes
 * ntube independent 1D problems, each running the cans1d main loop
 * (cans1d/md_mhdshktb/hdmlw/main.f:132-186) to its own end:
 *     ns=ns+1; cfl_m (common/cfl_m.f:2); stop if merr; t=t+dt;
 *     mlw_m (hdmlw/mlw_m.f:2); bdfrex low and high on ro,pr,vx,vy,by
 *     (bc/bdfrex.f:30-40); repeat while ns < nstop and t < tend
 * with t starting at 0.  Packed layout: tube k is cells off[k] .. off[k]+ix[k]-1
 * of every array (ix[k] = nx_k + 2*margin, ghosts included); off, ix, order
 * and the outputs t_out, ns_out, merr_out are device arrays of ntube entries.
 * order (NULL: 0..ntube-1) is the queue order of the persistent kernel --
 * longest tubes first balances the GPU.  The state ro..by is updated in
 * place.  mixed = 0 is bit-identical to the reference loop per tube;
 * mixed = 1 evaluates the cell and face fluxes in FP32 and keeps the
 * half-step state with its pressure recovery, the accumulation, the update
 * and cfl_m in FP64.  block_threads in {64,128,256,512}, blocks_per_cu
 * 1..32 (block_threads*blocks_per_cu <= 2048).  ws: cans1d_batch_ws_bytes
 * (tune = 1 adds room for cans1d_batch_tune's copy of the state).
 * No single counterpart in the reference, whose 1D driver runs one tube. */
size_t cans1d_batch_ws_bytes(long long ncells, int tune);
int cans1d_batch_run(int ntube, const long long *off, const int *ix, const int *order, double *ro, double *pr,
                     double *vx, double *vy, double *by, const double *bx, const double *bxm, const double *dx,
                     const double *dxm, long long ncells, double gm, double qav, double safety, double dtmin,
                     double tend, int nstop, int margin, int mixed, double *t_out, int *ns_out, int *merr_out,
                     int block_threads, int blocks_per_cu, void *ws, void *stream);
/* cans1d_batch_run with a step limit per tube: tube k runs while
 * ns < nstop_tube[k] and t < tend (nstop_tube: device array of ntube
 * entries).  With tend = +inf every tube takes exactly its own number of
 * steps -- e.g. the counts an FP64 run to tend took, so that a mixed run
 * can be compared with it at the same steps. */
int cans1d_batch_run_steps(int ntube, const long long *off, const int *ix, const int *order, double *ro,
                           double *pr, double *vx, double *vy, double *by, const double *bx, const double *bxm,
                           const double *dx, const double *dxm, long long ncells, double gm, double qav,
                           double safety, double dtmin, double tend, const int *nstop_tube, int margin, int mixed,
                           double *t_out, int *ns_out, int *merr_out, int block_threads, int blocks_per_cu,
                           void *ws, void *stream);
/* Batched periodic 2D runs: nrun independent problems, each running the
 * cans2d main loop (cans2d/md_mhdshktb/hdmlw/main.f:150-206 with bdperx and
 * bdpery at ghost width `margin`):
 *     ns=ns+1; cfl_m (cans2d/common/cfl_m.f:2-3); stop if merr; t=t+dt;
 *     mlw_m (cans2d/hdmlw/mlw_m.f:2-3); bdperx; bdpery; repeat while
 *     ns < nstop and t < tend
 * to its own end, in one persistent kernel (a workgroup per run at a time,
 * queue order `order`, NULL: 0..nrun-1).  Run k: fields ro,pr,vx,vy,bx,by,az
 * (state7: 7 device pointers) at cells off[k] .. off[k]+ix[k]*jx[k]-1,
 * column-major (ix,jx) with the ghost layers (filled periodically on entry);
 * its dx, dxm at entries offx[k] .. offx[k]+ix[k]-1 of the packed x arrays
 * (nx1d entries), dy, dym at offy[k] .. of the packed y arrays (ny1d).  off,
 * ix, jx, offx, offy, order, t_out, ns_out, merr_out are device arrays of
 * nrun entries.  FP64, bit-identical per run to the reference loop.  Every
 * call starts each run's loop at t = 0, ns = 0 from the state it finds and
 * leaves the final state in state7, t, ns and merr (9001 if dt < dtmin).
 * block_threads in {64,...,512} with block_threads*blocks_per_cu <= 512.
 * ws: cans2d_batch_ws_bytes (tune = 1 adds room for cans2d_batch_tune's
 * copy of the state).  The call synchronises the stream once (to upload the
 * per-run argument table into ws), then launches asynchronously.  No single
 * counterpart in the reference, whose 2D driver runs one problem. */
size_t cans2d_batch_ws_bytes(long long ncells, long long nx1d, long long ny1d, int nrun, int tune);
int cans2d_batch_run(int nrun, const long long *off, const int *ix, const int *jx, const long long *offx,
                     const long long *offy, const int *order, double *const *state7, const double *dx,
                     const double *dxm, const double *dy, const double *dym, long long ncells, long long nx1d,
                     long long ny1d, double gm, double qav, double safety, double dtmin, double tend, int nstop,
                     int margin, double *t_out, int *ns_out, int *merr_out, int block_threads, int blocks_per_cu,
                     void *ws, void *stream);
/* Launch-shape autotuner of cans2d_batch_run: times every (block_threads,
 * blocks_per_cu) with block_threads*blocks_per_cu <= 512 on a copy of the
 * state (probe_steps steps of every run; the state is unchanged, t_out,
 * ns_out, merr_out receive the probe's values) and returns the fastest;
 * cached per (nrun, ncells, nx1d+ny1d, probe_steps).  ws must have been
 * sized with tune = 1. */
int cans2d_batch_tune(int nrun, const long long *off, const int *ix, const int *jx, const long long *offx,
                      const long long *offy, const int *order, double *const *state7, const double *dx,
                      const double *dxm, const double *dy, const double *dym, long long ncells, long long nx1d,
                      long long ny1d, double gm, double qav, double safety, double dtmin, int margin, int probe_steps,
                      double *t_out, int *ns_out, int *merr_out, void *ws, int *block_threads, int *blocks_per_cu,
                      float *best_ms, void *stream);
/* cans2d_batch_run captured as an instantiated hipGraph (*graph; replay
 * with cans_graph_launch, free with cans_graph_destroy).  init7 (7 device
 * pointers, or NULL): copied into state7 at the start of every replay, so
 * replays repeat the same run.  The per-run argument table is uploaded into
 * ws before the capture; the graph reads it from there, so ws must not be
 * used by other calls while the graph is in use. */
int cans2d_batch_graph(int nrun, const long long *off, const int *ix, const int *jx, const long long *offx,
                       const long long *offy, const int *order, double *const *state7, const double *dx,
                       const double *dxm, const double *dy, const double *dym, long long ncells, long long nx1d,
                       long long ny1d, double gm, double qav, double safety, double dtmin, double tend, int nstop,
                       int margin, double *t_out, int *ns_out, int *merr_out, int block_threads, int blocks_per_cu,
                       void *ws, const double *const *init7, void **graph);
/* Autotuner of the persistent kernel's launch shape: times every
 * (block_threads, blocks_per_cu) on a copy of the state for probe_steps
 * steps per tube and returns the fastest (cached per batch shape for the
 * process).  The state is not changed; t_out/ns_out/merr_out are. */
int cans1d_batch_tune(int ntube, const long long *off, const int *ix, const int *order, double *ro, double *pr,
                      double *vx, double *vy, double *by, const double *bx, const double *bxm, const double *dx,
                      const double *dxm, long long ncells, double gm, double qav, double safety, double dtmin,
                      int margin, int mixed, int probe_steps, double *t_out, int *ns_out, int *merr_out, void *ws,
                      int *block_threads, int *blocks_per_cu, float *best_ms, void *stream);
/* The whole batch run as an instantiated hipGraph (optionally preceded by a
 * copy of init5[0..4] into ro..by, so every replay is the same run). */
int cans1d_batch_graph(int ntube, const long long *off, const int *ix, const int *order, double *ro, double *pr,
                       double *vx, double *vy, double *by, const double *bx, const double *bxm, const double *dx,
                       const double *dxm, long long ncells, double gm, double qav, double safety, double dtmin,
                       double tend, int nstop, int margin, int mixed, double *t_out, int *ns_out, int *merr_out,
                       int block_threads, int blocks_per_cu, void *ws, const double *const *init5, void **graph);
int cans_graph_launch(void *graph, void *stream);
int cans_graph_destroy(void *graph);

/* ---------------------------------------------------------------- UTRIP spicule model
 * cans1d/hdmlw/mlw_m_bg.f:2-3 mlw_m_bg(ro,pr,vx,vy,by,bx,bxm,dt,qav,gm,gx,gxm,sc,dsc,
 * scm,dscm,rr,rrm,drr,drrm,dx,dxm,ix): MLW for 1D MHD in a flux tube of
 * cross-section sc (face values scm), curvature radius rr, gravity gx (fuses
 * mlwhalf.f, mlwfull.f, mlwartv.f, mlwsrch.f, mlwsrcf.f); device arrays,
 * updated in place on i=2..ix-1. */
int cans1d_mlw_m_bg(double *ro, double *pr, double *vx, double *vy, double *by, const double *bx,
                    const double *bxm, double dt, double qav, double gm, const double *gx, const double *gxm,
                    const double *sc, const double *dsc, const double *scm, const double *dscm, const double *rr,
                    const double *rrm, const double *drr, const double *drrm, const double *dx, const double *dxm,
                    int ix, void *stream);
/* cans1d/md_spicule/hdmlw/trqalfven.f:2 trqalfven(vy,dt,time,x,ix,rr):
 * vy(i) += dt*30*q(i)*trq(i), q(i) = 2 rr(i) (anum_i - 0.5), q(ix) = q(ix-2).
 * trq is the time-independent profile tanh((x-0.75)/0.075)-1 with
 * trq(ix) = trq(ix-2), computed by the caller (host libm tanh) once.  The
 * noise anum_i is the i-th number of the Park-Miller stream of
 * cans1d/common/rangen.f from *mseed (host, 1..2^31-2), in place of the
 * compiler's RANDOM_NUMBER; *mseed is advanced by ix draws. */
int cans1d_trqalfven(double *vy, double dt, const double *trq, const double *rr, int ix, int *mseed,
                     void *stream);
/* cans1d/common/chkdav.f:2 chkdav(n_events,da,vx,floor,ix): da < floor ->
 * da = floor, vx = 0.  n_events is a host pointer (the call synchronises). */
int cans1d_chkdav(int *n_events, double *da, double *vx, double floor_, int ix, void *stream);
/* The time loop of cans1d/md_spicule/hdmlw/main.f:169-227 for nrun
 * independent realizations, all on the device in one persistent kernel:
 *     ns=ns+1; cfl_m (stop with merr=9001 if dt<dtmin); t=t+dt; mlw_m_bg;
 *     trqalfven; bnd (bdsppx/bdspnx low, bdfrex high: md_spicule/hdmlw/bnd.f);
 *     chkdav(ro); chkdav(pr); repeat while ns < nstop and t < tend
 * Realization r: fields state5[0..4] (ro,pr,vx,vy,by) at r*ix, updated in
 * place; geometry geo14 = bx,bxm,gx,gxm,sc,dsc,scm,dscm,rr,rrm,drr,drrm,dx,dxm
 * and the torque profile trq at r*ix if per_run_geometry, else shared;
 * t[r], ns[r] (loop entries so far), seed[r] (Park-Miller, 1..2^31-2) are
 * read and written back; merr[r] is written.  Calling again continues the
 * runs bit for bit.  snap (NULL: none) receives the state whenever
 * int(t/dtout) changes, as main.f writes its DAC output: snap[(r*nsnap_max+k)*5*ix
 * + q*ix + i], snap_t[r*nsnap_max+k], nsnap[r] of them (at most nsnap_max).
 * block_threads in {64,128,256,512,1024}; blocks <= 0: one per realization.
 * ws: cans1d_spicule_ws_bytes.  FP64, bit-identical to the reference loop
 * with the rangen noise.  The reference runs one realization per program. */
size_t cans1d_spicule_ws_bytes(int ix, int nrun, int per_run_geometry);
int cans1d_spicule_run(int nrun, int ix, int margin, double *const *state5, const double *const *geo14,
                       int per_run_geometry, const double *trq, double gm, double qav, double safety, double dtmin,
                       double floor_, double tend, int nstop, double *t, int *ns, int *merr, int *seed,
                       double dtout, double *snap, double *snap_t, int nsnap_max, int *nsnap, int block_threads,
                       int blocks, void *ws, void *stream);

/* HIP events for timing the engine kernels of a run: `events` of the run
 * entry points may be NULL or hold 2*nsteps events; the engine kernel of the
 * s-th step of the call is bracketed by events[2s] and events[2s+1], recorded
 * on the run's stream. */
int cans_event_create(void **ev);
int cans_event_destroy(void *ev);
int cans_event_record(void *ev, void *stream);
int cans_event_elapsed_ms(void *start, void *stop, float *ms);

/* ---------------------------------------------------------------- host arrays
 * Device mirrors of host arrays, used by the Fortran drop-in wrappers.  In the
 * default (strict) mode every wrapper call copies its arrays in and out; in
 * resident mode (cans_gpu_set_resident(1)) arrays stay on the device,
 * cans_gpu_unmap(copy_out = 1) marks a mirror changed, and
 * cans_gpu_sync_host() copies back the changed mirrors only (call it before
 * output; leaving resident mode does it too).  A program that changes a
 * mapped array on the host calls cans_gpu_invalidate on it; one that
 * releases a mapped array calls cans_gpu_forget on it first, so that a later
 * sync never writes to the released memory and a new array at the same
 * address is copied in afresh.
 * Workspace of the device-array entry points: one pool per (device, stream),
 * grown to the largest call and kept until cans_gpu_release_all; calls on
 * different streams never share one. */
void *cans_gpu_map(const void *host, size_t bytes, int copy_in);
int cans_gpu_unmap(void *host, int copy_out);
int cans_gpu_set_resident(int on);
int cans_gpu_sync_host(void);
int cans_gpu_invalidate(const void *host);
int cans_gpu_forget(const void *host);
int cans_gpu_release_all(void);

#ifdef __cplusplus
}
#endif
#endif
