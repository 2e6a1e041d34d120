/*
 * cans_oracle.h -- CPU restatement of the CANS MHD time step (test
 * infrastructure only).
 *
 * This is the checker the GPU kernels are compared against.  It restates the
 * reference Fortran routines in plain C with the reference's arithmetic order,
 * its literal kinds (REAL*4 constants such as 1.0e-4 and 1.e20 promoted to
 * double) and its MAX/MIN/SIGN semantics, and it keeps the reference's
 * array-at-a-time structure (one full sweep per reference loop) so that it is
 * a different program from the fused GPU kernels.  Each function cites the
 * reference file:line it follows (paths relative to cans_beta_vary/).
 *
 * Conventions (same as the reference): arrays are Fortran column-major,
 * dimension (ix[,jx[,kx]]) including `margin` ghost layers; index i of the
 * reference is element i-1 here.  All arithmetic is IEEE binary64 with no
 * FMA contraction (build with -ffp-contract=off).
 *
 * Only tests/, the smoke check and the CPU-baseline leg of bench.py may load
 * this library.  It is never part of the product path.
 */
#ifndef CANS_ORACLE_H
#define CANS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- 1D (cans1d) ---------------- */
void or1d_mlw_m(double *ro, double *pr, double *vx, double *vy, double *by,
                const double *bx, const double *bxm, double dt, double qav,
                double gm, const double *dx, const double *dxm, int ix);
void or1d_roe_m(double *ro, double *pr, double *vx, double *vy, double *by,
                const double *bx, const double *bxm, double dt, double gm,
                const double *dx, int ix);
void or1d_cfl_m(double *dt, double safety, double dtmin, int *merr, double gm,
                const double *bx, const double *ro, const double *pr,
                const double *vx, const double *vy, const double *by,
                const double *dx, int ix, int *imin);
/* kind: 0 fre, 1 spp, 2 spn, 3 smp, 4 smn, 5 cns (q0), 6 frd (dxm) */
void or1d_bd(int kind, int mbnd, int margin, double *qq, double q0,
             const double *dxm, int ix);
void or1d_bdperx(int margin0, int margin1, double *qq, int ix);
void or1d_grdrdy(double *dx, double *xm, double *dxm, const double *x, int ix);
void or1d_mlw_m_bg(double *ro, double *pr, double *vx, double *vy, double *by,
                   const double *bx, const double *bxm, double dt, double qav,
                   double gm, const double *gx, const double *gxm,
                   const double *sc, const double *dsc, const double *scm,
                   const double *dscm, const double *rr, const double *rrm,
                   const double *drr, const double *drrm, const double *dx,
                   const double *dxm, int ix);
void or1d_roe_m3(double *ro, double *pr, double *vx, double *vy, double *vz,
                 double *by, double *bz, const double *bx, const double *bxm,
                 double dt, double gm, const double *dx, int ix);
void or1d_mlw_m3(double *ro, double *pr, double *vx, double *vy, double *vz,
                 double *by, double *bz, const double *bx, const double *bxm,
                 double dt, double qav, double gm, const double *dx,
                 const double *dxm, int ix);
void or1d_cfl_m3(double *dt, double safety, double dtmin, int *merr, double gm,
                 const double *bx, const double *ro, const double *pr,
                 const double *vx, const double *vy, const double *by,
                 const double *vz, const double *bz, const double *dx, int ix,
                 int *imin);
void or1d_scrdy(double *dsc, double *dscm, const double *sc, const double *scm,
                const double *dx, const double *dxm, int ix);
void or1d_chkdav(int *n_events, double *da, double *vx, double floor_, int ix);
double or_rangen(int *mseed);

/* ---------------- UTRIP spicule loop (cans1d/md_spicule/hdmlw) ---------------- */
void orsp_trqalfven(double *vy, double dt, double time, const double *x, int ix,
                    const double *rr, int *mseed);
void orsp_bnd(int margin, double *ro, double *pr, double *vx, double *vy,
              double *by, int ix);
void orsp_bndsc(int margin, double *sc, double *dsc, double *scm, double *dscm,
                int ix);
void orsp_run(double *ro, double *pr, double *vx, double *vy, double *by,
              const double *const *g, int ix, int margin, double gm,
              double qav, double safety, double dtmin, double floor_,
              double tend, int nstop, double *t, int *ns, int *merr,
              int *mseed);

/* ---------------- 2D (cans2d) ---------------- */
void or2d_mlw_m(double *ro, double *pr, double *vx, double *vy, double *bx,
                double *by, double *az, double dt, double qav, double gm,
                const double *dx, const double *dxm, int ix,
                const double *dy, const double *dym, int jx);
void or2d_roe_m(double *ro, double *pr, double *vx, double *vy, double *bx,
                double *by, double *az, double dt, double gm,
                const double *dx, int ix, const double *dy, int jx);
void or2d_cfl_m(double *dt, double safety, double dtmin, int *merr, double gm,
                const double *ro, const double *pr, const double *vx,
                const double *vy, const double *bx, const double *by,
                const double *dx, int ix, const double *dy, int jx,
                int *imin, int *jmin);
void or2d_roe_m3(double *ro, double *pr, double *vx, double *vy, double *vz,
                 double *bx, double *by, double *bz, double *az, double dt,
                 double gm, const double *dx, int ix, const double *dy, int jx);
void or2d_mlw_m_c(double *ro, double *pr, double *vx, double *vz, double *bx,
                  double *bz, double *ay, double dt, double qav, double gm,
                  const double *x, const double *xm, const double *dx,
                  const double *dxm, int ix, const double *dz,
                  const double *dzm, int jx);
void or2d_mlw_m_s(double *ro, double *pr, double *vx, double *vy, double *bx, double *by, double *az,
                  double dt, double qav, double gm, const double *x, const double *xm, const double *dx,
                  const double *dxm, int ix, const double *y, const double *ym, const double *dy,
                  const double *dym, int jx);
void or2d_cfl_m_s(double *dt, double safety, double dtmin, int *merr, double gm,
                  const double *ro, const double *pr, const double *vx, const double *vy, const double *bx,
                  const double *by, const double *x, const double *dx, int ix, const double *dy, int jx,
                  int *imin, int *jmin);
void or2d_mlw_m3(double *ro, double *pr, double *vx, double *vy, double *vz,
                 double *bx, double *by, double *bz, double *az, double dt,
                 double qav, double gm, const double *dx, const double *dxm,
                 int ix, const double *dy, const double *dym, int jx);
void or2d_cfl_m3(double *dt, double safety, double dtmin, int *merr, double gm,
                 const double *ro, const double *pr, const double *vx,
                 const double *vy, const double *vz, const double *bx,
                 const double *by, const double *bz, const double *dx, int ix,
                 const double *dy, int jx, int *imin, int *jmin);
/* dir: 0 = x, 1 = y; kind as in or1d_bd, 7 = periodic (mbnd = margin0,
 * margin = margin1) */
void or2d_bd(int dir, int kind, int mbnd, int margin, double *qq, double q0,
             const double *dm, int ix, int jx);
void or2d_bbtoaa(double *az, const double *bx, const double *by,
                 const double *dxm, const double *dym, int margin, int ix,
                 int jx);

/* ---------------- 3D (cans3d) ---------------- */
void or3d_mlw_m(double *ro, double *pr, double *vx, double *vy, double *vz,
                double *bx, double *by, double *bz, double dt, double qav,
                double gm, const double *dx, const double *dxm, int ix,
                const double *dy, const double *dym, int jx,
                const double *dz, const double *dzm, int kx);
void or3d_mlw_m_c(double *ro, double *pr, double *vx, double *vy, double *vz,
                  double *bx, double *by, double *bz, double dt, double qav,
                  double gm, const double *x, const double *xm, const double *dx, const double *dxm, int ix,
                  const double *dy, const double *dym, int jx,
                  const double *dz, const double *dzm, int kx);
void or3d_mlw_m_s(double *ro, double *pr, double *vx, double *vy, double *vz,
                  double *bx, double *by, double *bz, double dt, double qav,
                  double gm, const double *x, const double *xm, const double *dx, const double *dxm, int ix,
                  const double *y, const double *ym, const double *dy, const double *dym, int jx,
                  const double *dz, const double *dzm, int kx);
void or3d_roe_m(double *ro, double *pr, double *vx, double *vy, double *vz,
                double *bx, double *by, double *bz, double dt, double gm,
                const double *dx, int ix, const double *dy, int jx,
                const double *dz, int kx);
void or3d_cfl_m(double *dt, double safety, double dtmin, int *merr, double gm,
                const double *ro, const double *pr, const double *vx,
                const double *vy, const double *vz, const double *bx,
                const double *by, const double *bz, const double *dx, int ix,
                const double *dy, int jx, const double *dz, int kx,
                int *imin, int *jmin, int *kmin);
void or3d_cfl_m_s(double *dt, double safety, double dtmin, int *merr, double gm,
                  const double *ro, const double *pr, const double *vx,
                  const double *vy, const double *vz, const double *bx,
                  const double *by, const double *bz, const double *x, const double *dx, int ix,
                  const double *y, const double *dy, int jx, const double *dz, int kx,
                  int *imin, int *jmin, int *kmin);
/* dir: 0 x, 1 y, 2 z; kind as or2d_bd */
void or3d_bd(int dir, int kind, int mbnd, int margin, double *qq, double q0,
             const double *dm, int ix, int jx, int kx);

/* ------------- HLLD MUSCL engine (cansxx-nosupport, mu units) ------------- */
void orh_engine(double *ro, double *pr, double *vx, double *vy, double *vz,
                double *bx, double *by, double *bz, double gm, double mu,
                double dt, const double *dx, const double *dy,
                const double *dz, int ix, int jx, int kx, const int *mfdim);

/* ---------------- helper routines with the reference argument lists
 * (hdmlw/mlwhalf.f, mlwfull.f, mlwartv.f, mlwsrch.f, mlwsrcf.f of each
 * dimensionality; hdroe/roeflux_m.f, tvdminmod.f; cansxx-nosupport
 * hdhlld/hlld_flux.f90, hlld_mh.f).  w arrays and daw carry a trailing
 * dimension 2 (left state first). */
void or1d_mlwhalf(const double *u, double *un, double *du, double dt, const double *f, const double *dxi,
                  const double *dxim, int ix);
void or1d_mlwfull(double *du, double dt, const double *f, const double *dxi, int ix);
void or1d_mlwartv(const double *u, double *du, double dt, const double *qx, const double *dxi, const double *dxim,
                  int ix);
void or1d_mlwsrch(double *un, double *du, double dt, const double *s, int ix);
void or1d_mlwsrcf(double *du, double dt, const double *s, const double *ux0, const double *ux1, int ix);
void or2d_mlwhalf(const double *u, double *un, double *du, double dt, const double *fx, const double *dxi,
                  const double *dxim, int ix, const double *fy, const double *dyi, const double *dyim, int jx);
void or2d_mlwfull(double *du, double dt, const double *fx, const double *dxi, const double *ux0, const double *ux1,
                  int ix, const double *fy, const double *dyi, const double *uy0, const double *uy1, int jx);
void or2d_mlwartv(const double *u, double *du, double dt, const double *qx, const double *dxi, const double *dxim,
                  int ix, const double *qy, const double *dyi, const double *dyim, int jx);
void or2d_mlwsrch(double *un, double *du, double dt, const double *s, int ix, int jx);
void or2d_mlwsrcf(double *du, double dt, const double *s, const double *ux0, const double *ux1, int ix,
                  const double *uy0, const double *uy1, int jx);
void or3d_mlwhalf(const double *u, double *un, double *du, double dt, const double *fx, const double *dxi,
                  const double *dxim, int ix, const double *fy, const double *dyi, const double *dyim, int jx,
                  const double *fz, const double *dzi, const double *dzim, int kx);
void or3d_mlwfull(double *du, double dt, const double *fx, const double *dxi, const double *ux0, const double *ux1,
                  int ix, const double *fy, const double *dyi, const double *uy0, const double *uy1, int jx,
                  const double *fz, const double *dzi, const double *uz0, const double *uz1, int kx);
void or3d_mlwartv(const double *u, double *du, double dt, const double *qx, const double *dxi, const double *dxim,
                  int ix, const double *qy, const double *dyi, const double *dyim, int jx, const double *qz,
                  const double *dzi, const double *dzim, int kx);
void or3d_mlwsrch(double *un, double *du, double dt, const double *s, int ix, int jx, int kx);
void or3d_mlwsrcf(double *du, double dt, const double *s, const double *ux0, const double *ux1, int ix,
                  const double *uy0, const double *uy1, int jx, const double *uz0, const double *uz1, int kx);
void or1d_roeflux_m(double *fro, double *fee, double *frx, double *fry, double *fby, double gm, const double *row,
                    const double *prw, const double *vxw, const double *vyw, const double *bxw, const double *byw,
                    int ix);
void or2d_roeflux_m(double *fro, double *fee, double *frx, double *fry, double *fby, double gm, const double *row,
                    const double *prw, const double *vxw, const double *vyw, const double *bxw, const double *byw,
                    int ix, int jx);
void or3d_roeflux_m(double *fro, double *fee, double *frx, double *fry, double *frz, double *fby, double *fbz,
                    double gm, const double *row, const double *prw, const double *vxw, const double *vyw,
                    const double *vzw, const double *bxw, const double *byw, const double *bzw, int ix, int jx,
                    int kx);
void or1d_tvdminmod(const double *da, double *daw, int ix);
void or2d_tvdminmod(int mdir, const double *da, double *daw, int ix, int jx);
void or3d_tvdminmod(int mdir, const double *da, double *daw, int ix, int jx, int kx);
void orh_hlld_flux(const double *row, const double *prw, const double *vxw, const double *vyw, const double *vzw,
                   const double *bx, const double *byw, const double *bzw, double gm, int ix, int jx, int kx,
                   double zfloor, double *fro, double *fee, double *frx, double *fry, double *frz, double *fby,
                   double *fbz);
void orh_hlld_mh(double *rof, double *prf, double *vxf, double *vyf, double *vzf, double *bxf, double *byf,
                 double *bzf, double *ron, double *een, double *rxn, double *ryn, double *rzn, double *bxn,
                 double *byn, double *bzn, double gm, double dt, const double *dx, const double *dy,
                 const double *dz, int ix, int jx, int kx, const int *mfdim, int mstage);
/* the cans1d main.f loop over routines passed as function pointers
 * (cans_oracle_refloop.c) */
int orref_loop_1d(void *cfl, void *mlw, void *bdfre, double *ro, double *pr, double *vx, double *vy, double *by,
                  const double *bx, const double *bxm, const double *dx, const double *dxm, int ix, int margin,
                  double gm, double safety, double qav, double dtmin, double tend, int nstop, double *t_out,
                  int *merr_out);

#ifdef __cplusplus
}
#endif
#endif
