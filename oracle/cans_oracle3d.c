/*
 * cans_oracle3d.c -- C restatement of the cans3d MLW engine, cfl_m and bc
 * (test infrastructure only; see cans_oracle.h).  Citations are to
 * cans_beta_vary/cans3d/.
 */
#include "cans_oracle.h"
#include "cans_oracle_util.h"

/* hdmlw/mlwhalf.f:38-68 */
static void mlwhalf3(const double *u, double *un, double *du, double dt,
                     const double *fx, const double *dxi, const double *dxim, int ix,
                     const double *fy, const double *dyi, const double *dyim, int jx,
                     const double *fz, const double *dzi, const double *dzim, int kx) {
    for (int k = 2; k <= kx - 1; k++)
        for (int j = 2; j <= jx - 1; j++)
            for (int i = 2; i <= ix - 1; i++)
                du[I3(i, j, k)] = du[I3(i, j, k)]
                    - 0.5 * dt * (0.5 * dxi[I1(i)] * (fx[I3(i + 1, j, k)] - fx[I3(i - 1, j, k)]))
                    - 0.5 * dt * (0.5 * dyi[I1(j)] * (fy[I3(i, j + 1, k)] - fy[I3(i, j - 1, k)]))
                    - 0.5 * dt * (0.5 * dzi[I1(k)] * (fz[I3(i, j, k + 1)] - fz[I3(i, j, k - 1)]));
    for (int k = 1; k <= kx - 1; k++)
        for (int j = 1; j <= jx - 1; j++)
            for (int i = 1; i <= ix - 1; i++) {
                double uh = (u[I3(i + 1, j, k)] + u[I3(i, j, k)] + u[I3(i + 1, j + 1, k)] + u[I3(i, j + 1, k)]
                             + u[I3(i + 1, j, k + 1)] + u[I3(i, j, k + 1)] + u[I3(i + 1, j + 1, k + 1)] + u[I3(i, j + 1, k + 1)]) / 8.0;
                double dfdx = (fx[I3(i + 1, j, k)] - fx[I3(i, j, k)] + fx[I3(i + 1, j + 1, k)] - fx[I3(i, j + 1, k)]
                               + fx[I3(i + 1, j, k + 1)] - fx[I3(i, j, k + 1)] + fx[I3(i + 1, j + 1, k + 1)] - fx[I3(i, j + 1, k + 1)])
                              / 4.0 * dxim[I1(i)];
                double dfdy = (fy[I3(i, j + 1, k)] - fy[I3(i, j, k)] + fy[I3(i + 1, j + 1, k)] - fy[I3(i + 1, j, k)]
                               + fy[I3(i, j + 1, k + 1)] - fy[I3(i, j, k + 1)] + fy[I3(i + 1, j + 1, k + 1)] - fy[I3(i + 1, j, k + 1)])
                              / 4.0 * dyim[I1(j)];
                double dfdz = (fz[I3(i, j, k + 1)] - fz[I3(i, j, k)] + fz[I3(i + 1, j, k + 1)] - fz[I3(i + 1, j, k)]
                               + fz[I3(i, j + 1, k + 1)] - fz[I3(i, j + 1, k)] + fz[I3(i + 1, j + 1, k + 1)] - fz[I3(i + 1, j + 1, k)])
                              / 4.0 * dzim[I1(k)];
                un[I3(i, j, k)] = uh - dt * dfdx - dt * dfdy - dt * dfdz;
            }
}

/* hdmlw/mlwfull.f:34-55 */
static void mlwfull3(double *du, double dt, const double *fx, const double *dxi, const double *ux0,
                     const double *ux1, int ix, const double *fy, const double *dyi, const double *uy0,
                     const double *uy1, int jx, const double *fz, const double *dzi, const double *uz0,
                     const double *uz1, int kx) {
    for (int k = 2; k <= kx - 1; k++)
        for (int j = 2; j <= jx - 1; j++)
            for (int i = 2; i <= ix - 1; i++) {
                double dfdx = dxi[I1(i)] *
                    (uz0[I1(k)] * (uy0[I1(j)] * (fx[I3(i, j - 1, k - 1)] - fx[I3(i - 1, j - 1, k - 1)])
                                   + uy1[I1(j)] * (fx[I3(i, j, k - 1)] - fx[I3(i - 1, j, k - 1)]))
                     + uz1[I1(k)] * (uy0[I1(j)] * (fx[I3(i, j - 1, k)] - fx[I3(i - 1, j - 1, k)])
                                     + uy1[I1(j)] * (fx[I3(i, j, k)] - fx[I3(i - 1, j, k)])));
                double dfdy = dyi[I1(j)] *
                    (ux0[I1(i)] * (uz0[I1(k)] * (fy[I3(i - 1, j, k - 1)] - fy[I3(i - 1, j - 1, k - 1)])
                                   + uz1[I1(k)] * (fy[I3(i - 1, j, k)] - fy[I3(i - 1, j - 1, k)]))
                     + ux1[I1(i)] * (uz0[I1(k)] * (fy[I3(i, j, k - 1)] - fy[I3(i, j - 1, k - 1)])
                                     + uz1[I1(k)] * (fy[I3(i, j, k)] - fy[I3(i, j - 1, k)])));
                double dfdz = dzi[I1(k)] *
                    (uy0[I1(j)] * (ux0[I1(i)] * (fz[I3(i - 1, j - 1, k)] - fz[I3(i - 1, j - 1, k - 1)])
                                   + ux1[I1(i)] * (fz[I3(i, j - 1, k)] - fz[I3(i, j - 1, k - 1)]))
                     + uy1[I1(j)] * (ux0[I1(i)] * (fz[I3(i - 1, j, k)] - fz[I3(i - 1, j, k - 1)])
                                     + ux1[I1(i)] * (fz[I3(i, j, k)] - fz[I3(i, j, k - 1)])));
                du[I3(i, j, k)] = du[I3(i, j, k)] - 0.5 * dt * dfdx - 0.5 * dt * dfdy - 0.5 * dt * dfdz;
            }
}

/* hdmlw/mlwartv.f:36-48 */
static void mlwartv3(const double *u, double *du, double dt, const double *qx, const double *dxi,
                     const double *dxim, int ix, const double *qy, const double *dyi, const double *dyim,
                     int jx, const double *qz, const double *dzi, const double *dzim, int kx) {
    for (int k = 2; k <= kx - 1; k++)
        for (int j = 2; j <= jx - 1; j++)
            for (int i = 2; i <= ix - 1; i++) {
                size_t p = I3(i, j, k);
                du[p] = du[p]
                    + dt * (dxi[I1(i)] * (qx[p] * dxim[I1(i)] * (u[I3(i + 1, j, k)] - u[p])
                                          - qx[I3(i - 1, j, k)] * dxim[I1(i - 1)] * (u[p] - u[I3(i - 1, j, k)])))
                    + dt * (dyi[I1(j)] * (qy[p] * dyim[I1(j)] * (u[I3(i, j + 1, k)] - u[p])
                                          - qy[I3(i, j - 1, k)] * dyim[I1(j - 1)] * (u[p] - u[I3(i, j - 1, k)])))
                    + dt * (dzi[I1(k)] * (qz[p] * dzim[I1(k)] * (u[I3(i, j, k + 1)] - u[p])
                                          - qz[I3(i, j, k - 1)] * dzim[I1(k - 1)] * (u[p] - u[I3(i, j, k - 1)])));
            }
}

/* hdmlw/mlw_m.f:2-453 (3D MHD, 8 fields) */
void or3d_mlw_m(double *ro, double *pr, double *vx, double *vy, double *vz,
                double *bx, double *by, double *bz, double dt, double qav,
                double gm, const double *dx, const double *dxm, int ix,
                const double *dy, const double *dym, int jx,
                const double *dz, const double *dzm, int kx) {
    const double pi = acos(-1.0);
    const double pi8i = 1.0 / pi / 8.0;
    const double pi4i = 1.0 / pi / 4.0;
    size_t n = (size_t)ix * jx * kx;
    double *dxi = dalloc(ix), *dxim = dalloc(ix), *ux0 = dalloc(ix), *ux1 = dalloc(ix);
    double *dyi = dalloc(jx), *dyim = dalloc(jx), *uy0 = dalloc(jx), *uy1 = dalloc(jx);
    double *dzi = dalloc(kx), *dzim = dalloc(kx), *uz0 = dalloc(kx), *uz1 = dalloc(kx);
    double *ex = dalloc(n), *ey = dalloc(n), *ez = dalloc(n);
    double *ee = dalloc(n), *rx = dalloc(n), *ry = dalloc(n), *rz = dalloc(n);
    double *roh = dalloc(n), *eeh = dalloc(n), *rxh = dalloc(n), *ryh = dalloc(n), *rzh = dalloc(n);
    double *bxh = dalloc(n), *byh = dalloc(n), *bzh = dalloc(n);
    double *exh = dalloc(n), *eyh = dalloc(n), *ezh = dalloc(n);
    double *prh = dalloc(n), *vxh = dalloc(n), *vyh = dalloc(n), *vzh = dalloc(n);
    double *dro = dalloc(n), *dee = dalloc(n), *drx = dalloc(n), *dry = dalloc(n), *drz = dalloc(n);
    double *dbx = dalloc(n), *dby = dalloc(n), *dbz = dalloc(n);
    double *fx = dalloc(n), *fy = dalloc(n), *fz = dalloc(n), *qx = dalloc(n), *qy = dalloc(n), *qz = dalloc(n);

    for (int i = 1; i <= ix; i++) { dxi[I1(i)] = 1.0 / dx[I1(i)]; dxim[I1(i)] = 1.0 / dxm[I1(i)]; }
    for (int i = 2; i <= ix - 1; i++) { ux1[I1(i)] = 0.5 * dxm[I1(i - 1)] / dx[I1(i)]; ux0[I1(i)] = 0.5 * dxm[I1(i)] / dx[I1(i)]; }
    for (int j = 1; j <= jx; j++) { dyi[I1(j)] = 1.0 / dy[I1(j)]; dyim[I1(j)] = 1.0 / dym[I1(j)]; }
    for (int j = 2; j <= jx - 1; j++) { uy1[I1(j)] = 0.5 * dym[I1(j - 1)] / dy[I1(j)]; uy0[I1(j)] = 0.5 * dym[I1(j)] / dy[I1(j)]; }
    for (int k = 1; k <= kx; k++) { dzi[I1(k)] = 1.0 / dz[I1(k)]; dzim[I1(k)] = 1.0 / dzm[I1(k)]; }
    for (int k = 2; k <= kx - 1; k++) { uz1[I1(k)] = 0.5 * dzm[I1(k - 1)] / dz[I1(k)]; uz0[I1(k)] = 0.5 * dzm[I1(k)] / dz[I1(k)]; }
    for (size_t p = 0; p < n; p++) {
        dro[p] = dee[p] = drx[p] = dry[p] = drz[p] = dbx[p] = dby[p] = dbz[p] = 0.0;
    }
    for (size_t p = 0; p < n; p++) { /* :70-81 */
        double vv = sq(vx[p]) + sq(vy[p]) + sq(vz[p]);
        double bb = sq(bx[p]) + sq(by[p]) + sq(bz[p]);
        ee[p] = pr[p] / (gm - 1.0) + 0.5 * ro[p] * vv + pi8i * bb;
        rx[p] = ro[p] * vx[p];
        ry[p] = ro[p] * vy[p];
        rz[p] = ro[p] * vz[p];
    }
    for (size_t p = 0; p < n; p++) { /* :82-90 */
        ex[p] = -vy[p] * bz[p] + vz[p] * by[p];
        ey[p] = -vz[p] * bx[p] + vx[p] * bz[p];
        ez[p] = -vx[p] * by[p] + vy[p] * bx[p];
    }
#define HALF(u, un, du) mlwhalf3(u, un, du, dt, fx, dxi, dxim, ix, fy, dyi, dyim, jx, fz, dzi, dzim, kx)
    for (size_t p = 0; p < n; p++) { fx[p] = ro[p] * vx[p]; fy[p] = ro[p] * vy[p]; fz[p] = ro[p] * vz[p]; }
    HALF(ro, roh, dro);
    for (size_t p = 0; p < n; p++) {
        double vv = sq(vx[p]) + sq(vy[p]) + sq(vz[p]);
        double ep = pr[p] * gm / (gm - 1.0) + 0.5 * ro[p] * vv;
        fx[p] = ep * vx[p] + (bz[p] * ey[p] - by[p] * ez[p]) * pi4i;
        fy[p] = ep * vy[p] + (bx[p] * ez[p] - bz[p] * ex[p]) * pi4i;
        fz[p] = ep * vz[p] + (by[p] * ex[p] - bx[p] * ey[p]) * pi4i;
    }
    HALF(ee, eeh, dee);
    for (size_t p = 0; p < n; p++) {
        fx[p] = ro[p] * sq(vx[p]) + pr[p] + pi8i * (sq(by[p]) + sq(bz[p]) - sq(bx[p]));
        fy[p] = ro[p] * vx[p] * vy[p] - pi4i * bx[p] * by[p];
        fz[p] = ro[p] * vx[p] * vz[p] - pi4i * bx[p] * bz[p];
    }
    HALF(rx, rxh, drx);
    for (size_t p = 0; p < n; p++) {
        fx[p] = ro[p] * vy[p] * vx[p] - pi4i * by[p] * bx[p];
        fy[p] = ro[p] * sq(vy[p]) + pr[p] + pi8i * (sq(bx[p]) + sq(bz[p]) - sq(by[p]));
        fz[p] = ro[p] * vy[p] * vz[p] - pi4i * by[p] * bz[p];
    }
    HALF(ry, ryh, dry);
    for (size_t p = 0; p < n; p++) {
        fx[p] = ro[p] * vz[p] * vx[p] - pi4i * bz[p] * bx[p];
        fy[p] = ro[p] * vz[p] * vy[p] - pi4i * bz[p] * by[p];
        fz[p] = ro[p] * sq(vz[p]) + pr[p] + pi8i * (sq(bx[p]) + sq(by[p]) - sq(bz[p]));
    }
    HALF(rz, rzh, drz);
    for (size_t p = 0; p < n; p++) { fx[p] = 0.0; fy[p] = ez[p]; fz[p] = -ey[p]; }
    HALF(bx, bxh, dbx);
    for (size_t p = 0; p < n; p++) { fx[p] = -ez[p]; fy[p] = 0.0; fz[p] = ex[p]; }
    HALF(by, byh, dby);
    for (size_t p = 0; p < n; p++) { fx[p] = ey[p]; fy[p] = -ex[p]; fz[p] = 0.0; }
    HALF(bz, bzh, dbz);
#undef HALF
    for (int k = 1; k <= kx - 1; k++) /* :193-204 */
        for (int j = 1; j <= jx - 1; j++)
            for (int i = 1; i <= ix - 1; i++) {
                size_t p = I3(i, j, k);
                vxh[p] = rxh[p] / roh[p];
                vyh[p] = ryh[p] / roh[p];
                vzh[p] = rzh[p] / roh[p];
                double vv = sq(vxh[p]) + sq(vyh[p]) + sq(vzh[p]);
                double bb = sq(bxh[p]) + sq(byh[p]) + sq(bzh[p]);
                prh[p] = (gm - 1.0) * (eeh[p] - 0.5 * roh[p] * vv - pi8i * bb);
            }
    for (size_t p = 0; p < n; p++) {
        exh[p] = -vyh[p] * bzh[p] + vzh[p] * byh[p];
        eyh[p] = -vzh[p] * bxh[p] + vxh[p] * bzh[p];
        ezh[p] = -vxh[p] * byh[p] + vyh[p] * bxh[p];
    }
#define FULL(du) mlwfull3(du, dt, fx, dxi, ux0, ux1, ix, fy, dyi, uy0, uy1, jx, fz, dzi, uz0, uz1, kx)
#define HL for (int k = 1; k <= kx - 1; k++) for (int j = 1; j <= jx - 1; j++) for (int i = 1; i <= ix - 1; i++)
    HL { size_t p = I3(i, j, k); fx[p] = roh[p] * vxh[p]; fy[p] = roh[p] * vyh[p]; fz[p] = roh[p] * vzh[p]; }
    FULL(dro);
    HL {
        size_t p = I3(i, j, k);
        double vv = sq(vxh[p]) + sq(vyh[p]) + sq(vzh[p]);
        double ep = prh[p] * gm / (gm - 1.0) + 0.5 * roh[p] * vv;
        fx[p] = ep * vxh[p] + (bzh[p] * eyh[p] - byh[p] * ezh[p]) * pi4i;
        fy[p] = ep * vyh[p] + (bxh[p] * ezh[p] - bzh[p] * exh[p]) * pi4i;
        fz[p] = ep * vzh[p] + (byh[p] * exh[p] - bxh[p] * eyh[p]) * pi4i;
    }
    FULL(dee);
    HL {
        size_t p = I3(i, j, k);
        fx[p] = roh[p] * sq(vxh[p]) + prh[p] + pi8i * (sq(byh[p]) + sq(bzh[p]) - sq(bxh[p]));
        fy[p] = roh[p] * vxh[p] * vyh[p] - pi4i * bxh[p] * byh[p];
        fz[p] = roh[p] * vxh[p] * vzh[p] - pi4i * bxh[p] * bzh[p];
    }
    FULL(drx);
    HL {
        size_t p = I3(i, j, k);
        fx[p] = roh[p] * vyh[p] * vxh[p] - pi4i * byh[p] * bxh[p];
        fy[p] = roh[p] * sq(vyh[p]) + prh[p] + pi8i * (sq(bxh[p]) + sq(bzh[p]) - sq(byh[p]));
        fz[p] = roh[p] * vyh[p] * vzh[p] - pi4i * byh[p] * bzh[p];
    }
    FULL(dry);
    for (int k = 1; k <= kx - 1; k++) /* :269-280 runs j=1..jx, i=1..ix */
        for (int j = 1; j <= jx; j++)
            for (int i = 1; i <= ix; i++) {
                size_t p = I3(i, j, k);
                fx[p] = roh[p] * vzh[p] * vxh[p] - pi4i * bzh[p] * bxh[p];
                fy[p] = roh[p] * vzh[p] * vyh[p] - pi4i * bzh[p] * byh[p];
                fz[p] = roh[p] * sq(vzh[p]) + prh[p] + pi8i * (sq(bxh[p]) + sq(byh[p]) - sq(bzh[p]));
            }
    FULL(drz);
    HL { size_t p = I3(i, j, k); fx[p] = 0.0; fy[p] = ezh[p]; fz[p] = -eyh[p]; }
    FULL(dbx);
    HL { size_t p = I3(i, j, k); fx[p] = -ezh[p]; fy[p] = 0.0; fz[p] = exh[p]; }
    FULL(dby);
    HL { size_t p = I3(i, j, k); fx[p] = eyh[p]; fy[p] = -exh[p]; fz[p] = 0.0; }
    FULL(dbz);
#undef HL
#undef FULL
    const double zero = 0.0;
    for (int k = 1; k <= kx; k++)
        for (int j = 1; j <= jx; j++)
            for (int i = 1; i <= ix - 1; i++)
                qx[I3(i, j, k)] = qav * dxm[I1(i)] * fmaxF(zero, fabs(vx[I3(i + 1, j, k)] - vx[I3(i, j, k)]) - R4_1EM4);
    for (int k = 1; k <= kx; k++)
        for (int j = 1; j <= jx - 1; j++)
            for (int i = 1; i <= ix; i++)
                qy[I3(i, j, k)] = qav * dym[I1(j)] * fmaxF(zero, fabs(vy[I3(i, j + 1, k)] - vy[I3(i, j, k)]) - R4_1EM4);
    for (int k = 1; k <= kx - 1; k++)
        for (int j = 1; j <= jx; j++)
            for (int i = 1; i <= ix; i++)
                qz[I3(i, j, k)] = qav * dzm[I1(k)] * fmaxF(zero, fabs(vz[I3(i, j, k + 1)] - vz[I3(i, j, k)]) - R4_1EM4);
#define ARTV(u, du) mlwartv3(u, du, dt, qx, dxi, dxim, ix, qy, dyi, dyim, jx, qz, dzi, dzim, kx)
    ARTV(ro, dro); ARTV(ee, dee); ARTV(rx, drx); ARTV(ry, dry); ARTV(rz, drz);
    ARTV(bx, dbx); ARTV(by, dby); ARTV(bz, dbz);
#undef ARTV
    for (int k = 2; k <= kx - 1; k++)
        for (int j = 2; j <= jx - 1; j++)
            for (int i = 2; i <= ix - 1; i++) {
                size_t p = I3(i, j, k);
                ro[p] = ro[p] + dro[p]; ee[p] = ee[p] + dee[p]; rx[p] = rx[p] + drx[p]; ry[p] = ry[p] + dry[p];
                rz[p] = rz[p] + drz[p]; bx[p] = bx[p] + dbx[p]; by[p] = by[p] + dby[p]; bz[p] = bz[p] + dbz[p];
            }
    for (int k = 2; k <= kx - 1; k++)
        for (int j = 2; j <= jx - 1; j++)
            for (int i = 2; i <= ix - 1; i++) {
                size_t p = I3(i, j, k);
                vx[p] = rx[p] / ro[p];
                vy[p] = ry[p] / ro[p];
                vz[p] = rz[p] / ro[p];
                double vv = sq(vx[p]) + sq(vy[p]) + sq(vz[p]);
                double bb = sq(bx[p]) + sq(by[p]) + sq(bz[p]);
                pr[p] = (gm - 1.0) * (ee[p] - 0.5 * ro[p] * vv - pi8i * bb);
            }
    double *tofree[] = {dxi, dxim, ux0, ux1, dyi, dyim, uy0, uy1, dzi, dzim, uz0, uz1, ex, ey, ez, ee, rx, ry, rz,
                        roh, eeh, rxh, ryh, rzh, bxh, byh, bzh, exh, eyh, ezh, prh, vxh, vyh, vzh,
                        dro, dee, drx, dry, drz, dbx, dby, dbz, fx, fy, fz, qx, qy, qz};
    for (size_t t = 0; t < sizeof(tofree) / sizeof(tofree[0]); t++) free(tofree[t]);
}

/* hdmlw/mlwsrch.f:32-47 (3D): cell source into du, corner-averaged source into un */
static void mlwsrch3(double *un, double *du, double dt, const double *s, int ix, int jx, int kx) {
    for (int k = 2; k <= kx - 1; k++)
        for (int j = 2; j <= jx - 1; j++)
            for (int i = 2; i <= ix - 1; i++) du[I3(i, j, k)] = du[I3(i, j, k)] + 0.5 * dt * s[I3(i, j, k)];
    for (int k = 1; k <= kx - 1; k++)
        for (int j = 1; j <= jx - 1; j++)
            for (int i = 1; i <= ix - 1; i++) {
                double sh = (s[I3(i + 1, j, k)] + s[I3(i, j, k)] + s[I3(i + 1, j + 1, k)] + s[I3(i, j + 1, k)]
                             + s[I3(i + 1, j, k + 1)] + s[I3(i, j, k + 1)] + s[I3(i + 1, j + 1, k + 1)]
                             + s[I3(i, j + 1, k + 1)]) / 8.0;
                un[I3(i, j, k)] = un[I3(i, j, k)] + dt * sh;
            }
}

/* hdmlw/mlwsrcf.f:33-46 (3D): corner source, trilinear weights */
static void mlwsrcf3(double *du, double dt, const double *s, const double *ux0, const double *ux1, int ix,
                     const double *uy0, const double *uy1, int jx, const double *uz0, const double *uz1, int kx) {
    for (int k = 2; k <= kx - 1; k++)
        for (int j = 2; j <= jx - 1; j++)
            for (int i = 2; i <= ix - 1; i++) {
                double sh = uz0[I1(k)] * (uy0[I1(j)] * (ux0[I1(i)] * s[I3(i - 1, j - 1, k - 1)] + ux1[I1(i)] * s[I3(i, j - 1, k - 1)])
                                          + uy1[I1(j)] * (ux0[I1(i)] * s[I3(i - 1, j, k - 1)] + ux1[I1(i)] * s[I3(i, j, k - 1)]))
                          + uz1[I1(k)] * (uy0[I1(j)] * (ux0[I1(i)] * s[I3(i - 1, j - 1, k)] + ux1[I1(i)] * s[I3(i, j - 1, k)])
                                          + uy1[I1(j)] * (ux0[I1(i)] * s[I3(i - 1, j, k)] + ux1[I1(i)] * s[I3(i, j, k)]));
                du[I3(i, j, k)] = du[I3(i, j, k)] + 0.5 * dt * sh;
            }
}

/* hdmlw/mlw_m_c.f:2-489 (3D cylindrical (r, phi, z) MHD, 8 fields): mlw_m
 * with the phi fluxes divided by the radius (x at cells, xm at corners) and
 * the geometric sources of ro, ee, rx, ry, rz, bz through mlwsrch/mlwsrcf */
void or3d_mlw_m_c(double *ro, double *pr, double *vx, double *vy, double *vz,
                  double *bx, double *by, double *bz, double dt, double qav,
                  double gm, const double *x, const double *xm, const double *dx, const double *dxm, int ix,
                  const double *dy, const double *dym, int jx,
                  const double *dz, const double *dzm, int kx) {
    const double pi = acos(-1.0);
    const double pi8i = 1.0 / pi / 8.0;
    const double pi4i = 1.0 / pi / 4.0;
    size_t n = (size_t)ix * jx * kx;
    double *dxi = dalloc(ix), *dxim = dalloc(ix), *ux0 = dalloc(ix), *ux1 = dalloc(ix);
    double *dyi = dalloc(jx), *dyim = dalloc(jx), *uy0 = dalloc(jx), *uy1 = dalloc(jx);
    double *dzi = dalloc(kx), *dzim = dalloc(kx), *uz0 = dalloc(kx), *uz1 = dalloc(kx);
    double *ex = dalloc(n), *ey = dalloc(n), *ez = dalloc(n);
    double *ee = dalloc(n), *rx = dalloc(n), *ry = dalloc(n), *rz = dalloc(n);
    double *roh = dalloc(n), *eeh = dalloc(n), *rxh = dalloc(n), *ryh = dalloc(n), *rzh = dalloc(n);
    double *bxh = dalloc(n), *byh = dalloc(n), *bzh = dalloc(n);
    double *exh = dalloc(n), *eyh = dalloc(n), *ezh = dalloc(n);
    double *prh = dalloc(n), *vxh = dalloc(n), *vyh = dalloc(n), *vzh = dalloc(n);
    double *dro = dalloc(n), *dee = dalloc(n), *drx = dalloc(n), *dry = dalloc(n), *drz = dalloc(n);
    double *dbx = dalloc(n), *dby = dalloc(n), *dbz = dalloc(n);
    double *fx = dalloc(n), *fy = dalloc(n), *fz = dalloc(n), *ss = dalloc(n);
    double *qx = dalloc(n), *qy = dalloc(n), *qz = dalloc(n);

    for (int i = 1; i <= ix; i++) { dxi[I1(i)] = 1.0 / dx[I1(i)]; dxim[I1(i)] = 1.0 / dxm[I1(i)]; }
    for (int i = 2; i <= ix - 1; i++) { ux1[I1(i)] = 0.5 * dxm[I1(i - 1)] / dx[I1(i)]; ux0[I1(i)] = 0.5 * dxm[I1(i)] / dx[I1(i)]; }
    for (int j = 1; j <= jx; j++) { dyi[I1(j)] = 1.0 / dy[I1(j)]; dyim[I1(j)] = 1.0 / dym[I1(j)]; }
    for (int j = 2; j <= jx - 1; j++) { uy1[I1(j)] = 0.5 * dym[I1(j - 1)] / dy[I1(j)]; uy0[I1(j)] = 0.5 * dym[I1(j)] / dy[I1(j)]; }
    for (int k = 1; k <= kx; k++) { dzi[I1(k)] = 1.0 / dz[I1(k)]; dzim[I1(k)] = 1.0 / dzm[I1(k)]; }
    for (int k = 2; k <= kx - 1; k++) { uz1[I1(k)] = 0.5 * dzm[I1(k - 1)] / dz[I1(k)]; uz0[I1(k)] = 0.5 * dzm[I1(k)] / dz[I1(k)]; }
    for (size_t p = 0; p < n; p++) dro[p] = dee[p] = drx[p] = dry[p] = drz[p] = dbx[p] = dby[p] = dbz[p] = 0.0;
    for (size_t p = 0; p < n; p++) { /* :84-94 */
        double vv = sq(vx[p]) + sq(vy[p]) + sq(vz[p]);
        double bb = sq(bx[p]) + sq(by[p]) + sq(bz[p]);
        ee[p] = pr[p] / (gm - 1.0) + 0.5 * ro[p] * vv + pi8i * bb;
        rx[p] = ro[p] * vx[p];
        ry[p] = ro[p] * vy[p];
        rz[p] = ro[p] * vz[p];
    }
    for (size_t p = 0; p < n; p++) { /* :96-104 */
        ex[p] = -vy[p] * bz[p] + vz[p] * by[p];
        ey[p] = -vz[p] * bx[p] + vx[p] * bz[p];
        ez[p] = -vx[p] * by[p] + vy[p] * bx[p];
    }
#define CELLS for (int k = 1; k <= kx; k++) for (int j = 1; j <= jx; j++) for (int i = 1; i <= ix; i++)
#define HALF(u, un, du) mlwhalf3(u, un, du, dt, fx, dxi, dxim, ix, fy, dyi, dyim, jx, fz, dzi, dzim, kx)
    CELLS { /* :109-121 */
        size_t p = I3(i, j, k); double r = x[I1(i)];
        fx[p] = ro[p] * vx[p]; fy[p] = ro[p] * vy[p] / r; fz[p] = ro[p] * vz[p]; ss[p] = -(fx[p] / r);
    }
    HALF(ro, roh, dro); mlwsrch3(roh, dro, dt, ss, ix, jx, kx);
    CELLS { /* :125-142 */
        size_t p = I3(i, j, k); double r = x[I1(i)];
        double vv = sq(vx[p]) + sq(vy[p]) + sq(vz[p]);
        double ep = pr[p] * gm / (gm - 1.0) + 0.5 * ro[p] * vv;
        fx[p] = ep * vx[p] + (bz[p] * ey[p] - by[p] * ez[p]) * pi4i;
        fy[p] = ep * vy[p] + (bx[p] * ez[p] - bz[p] * ex[p]) * pi4i;
        fy[p] = fy[p] / r;
        fz[p] = ep * vz[p] + (by[p] * ex[p] - bx[p] * ey[p]) * pi4i;
        ss[p] = -(fx[p] / r);
    }
    HALF(ee, eeh, dee); mlwsrch3(eeh, dee, dt, ss, ix, jx, kx);
    CELLS { /* :145-162 */
        size_t p = I3(i, j, k); double r = x[I1(i)];
        fx[p] = ro[p] * sq(vx[p]) + pr[p] + pi8i * (sq(by[p]) + sq(bz[p]) - sq(bx[p]));
        fy[p] = ro[p] * vx[p] * vy[p] - pi4i * bx[p] * by[p];
        fy[p] = fy[p] / r;
        fz[p] = ro[p] * vx[p] * vz[p] - pi4i * bx[p] * bz[p];
        ss[p] = -(ro[p] / r * (sq(vx[p]) - sq(vy[p]))) + pi4i / r * (sq(bx[p]) - sq(by[p]));
    }
    HALF(rx, rxh, drx); mlwsrch3(rxh, drx, dt, ss, ix, jx, kx);
    CELLS { /* :165-182 */
        size_t p = I3(i, j, k); double r = x[I1(i)];
        fx[p] = ro[p] * vy[p] * vx[p] - pi4i * by[p] * bx[p];
        fy[p] = ro[p] * sq(vy[p]) + pr[p] + pi8i * (sq(bx[p]) + sq(bz[p]) - sq(by[p]));
        fy[p] = fy[p] / r;
        fz[p] = ro[p] * vy[p] * vz[p] - pi4i * by[p] * bz[p];
        ss[p] = -(2.0 * ro[p] / r * vx[p] * vy[p]) + 2.0 * pi4i / r * bx[p] * by[p];
    }
    HALF(ry, ryh, dry); mlwsrch3(ryh, dry, dt, ss, ix, jx, kx);
    CELLS { /* :185-201 */
        size_t p = I3(i, j, k); double r = x[I1(i)];
        fx[p] = ro[p] * vz[p] * vx[p] - pi4i * bz[p] * bx[p];
        fy[p] = ro[p] * vz[p] * vy[p] - pi4i * bz[p] * by[p];
        fy[p] = fy[p] / r;
        fz[p] = ro[p] * sq(vz[p]) + pr[p] + pi8i * (sq(bx[p]) + sq(by[p]) - sq(bz[p]));
        ss[p] = -(fx[p] / r);
    }
    HALF(rz, rzh, drz); mlwsrch3(rzh, drz, dt, ss, ix, jx, kx);
    CELLS { size_t p = I3(i, j, k); fx[p] = 0.0; fy[p] = ez[p] / x[I1(i)]; fz[p] = -ey[p]; } /* :205-213 */
    HALF(bx, bxh, dbx);
    CELLS { size_t p = I3(i, j, k); fx[p] = -ez[p]; fy[p] = 0.0; fz[p] = ex[p]; }          /* :218-226 */
    HALF(by, byh, dby);
    CELLS { /* :231-243 */
        size_t p = I3(i, j, k); double r = x[I1(i)];
        fx[p] = ey[p]; fy[p] = -ex[p] / r; fz[p] = 0.0; ss[p] = -(fx[p] / r);
    }
    HALF(bz, bzh, dbz); mlwsrch3(bzh, dbz, dt, ss, ix, jx, kx);
#undef HALF
#undef CELLS
    for (int k = 1; k <= kx - 1; k++) /* :248-258 */
        for (int j = 1; j <= jx - 1; j++)
            for (int i = 1; i <= ix - 1; i++) {
                size_t p = I3(i, j, k);
                vxh[p] = rxh[p] / roh[p];
                vyh[p] = ryh[p] / roh[p];
                vzh[p] = rzh[p] / roh[p];
                double vv = sq(vxh[p]) + sq(vyh[p]) + sq(vzh[p]);
                double bb = sq(bxh[p]) + sq(byh[p]) + sq(bzh[p]);
                prh[p] = (gm - 1.0) * (eeh[p] - 0.5 * roh[p] * vv - pi8i * bb);
            }
    for (int k = 1; k <= kx - 1; k++) /* :260-268 (exh..ezh used at corners 1..n-1 only) */
        for (int j = 1; j <= jx - 1; j++)
            for (int i = 1; i <= ix - 1; i++) {
                size_t p = I3(i, j, k);
                exh[p] = -vyh[p] * bzh[p] + vzh[p] * byh[p];
                eyh[p] = -vzh[p] * bxh[p] + vxh[p] * bzh[p];
                ezh[p] = -vxh[p] * byh[p] + vyh[p] * bxh[p];
            }
#define FULL(du) mlwfull3(du, dt, fx, dxi, ux0, ux1, ix, fy, dyi, uy0, uy1, jx, fz, dzi, uz0, uz1, kx)
#define SRCF(du) mlwsrcf3(du, dt, ss, ux0, ux1, ix, uy0, uy1, jx, uz0, uz1, kx)
#define HL for (int k = 1; k <= kx - 1; k++) for (int j = 1; j <= jx - 1; j++) for (int i = 1; i <= ix - 1; i++)
    HL { /* :273-287 */
        size_t p = I3(i, j, k); double r = xm[I1(i)];
        fx[p] = roh[p] * vxh[p]; fy[p] = roh[p] * vyh[p] / r; fz[p] = roh[p] * vzh[p]; ss[p] = -(fx[p] / r);
    }
    FULL(dro); SRCF(dro);
    HL { /* :290-308 */
        size_t p = I3(i, j, k); double r = xm[I1(i)];
        double vv = sq(vxh[p]) + sq(vyh[p]) + sq(vzh[p]);
        double ep = prh[p] * gm / (gm - 1.0) + 0.5 * roh[p] * vv;
        fx[p] = ep * vxh[p] + (bzh[p] * eyh[p] - byh[p] * ezh[p]) * pi4i;
        fy[p] = ep * vyh[p] + (bxh[p] * ezh[p] - bzh[p] * exh[p]) * pi4i;
        fy[p] = fy[p] / r;
        fz[p] = ep * vzh[p] + (byh[p] * exh[p] - bxh[p] * eyh[p]) * pi4i;
        ss[p] = -(fx[p] / r);
    }
    FULL(dee); SRCF(dee);
    HL { /* :311-328 */
        size_t p = I3(i, j, k); double r = xm[I1(i)];
        fx[p] = roh[p] * sq(vxh[p]) + prh[p] + pi8i * (sq(byh[p]) + sq(bzh[p]) - sq(bxh[p]));
        fy[p] = roh[p] * vxh[p] * vyh[p] - pi4i * bxh[p] * byh[p];
        fy[p] = fy[p] / r;
        fz[p] = roh[p] * vxh[p] * vzh[p] - pi4i * bxh[p] * bzh[p];
        ss[p] = -(roh[p] / r * (sq(vxh[p]) - sq(vyh[p]))) + pi4i / r * (sq(bxh[p]) - sq(byh[p]));
    }
    FULL(drx); SRCF(drx);
    HL { /* :331-348 */
        size_t p = I3(i, j, k); double r = xm[I1(i)];
        fx[p] = roh[p] * vyh[p] * vxh[p] - pi4i * byh[p] * bxh[p];
        fy[p] = roh[p] * sq(vyh[p]) + prh[p] + pi8i * (sq(bxh[p]) + sq(bzh[p]) - sq(byh[p]));
        fy[p] = fy[p] / r;
        fz[p] = roh[p] * vyh[p] * vzh[p] - pi4i * byh[p] * bzh[p];
        ss[p] = -(2.0 * roh[p] / r * vxh[p] * vyh[p]) + 2.0 * pi4i / r * bxh[p] * byh[p];
    }
    FULL(dry); SRCF(dry);
    HL { /* :351-367 */
        size_t p = I3(i, j, k); double r = xm[I1(i)];
        fx[p] = roh[p] * vzh[p] * vxh[p] - pi4i * bzh[p] * bxh[p];
        fy[p] = roh[p] * vzh[p] * vyh[p] - pi4i * bzh[p] * byh[p];
        fy[p] = fy[p] / r;
        fz[p] = roh[p] * sq(vzh[p]) + prh[p] + pi8i * (sq(bxh[p]) + sq(byh[p]) - sq(bzh[p]));
        ss[p] = -(fx[p] / r);
    }
    FULL(drz); SRCF(drz);
    HL { size_t p = I3(i, j, k); fx[p] = 0.0; fy[p] = ezh[p] / xm[I1(i)]; fz[p] = -eyh[p]; } /* :370-379 */
    FULL(dbx);
    HL { size_t p = I3(i, j, k); fx[p] = -ezh[p]; fy[p] = 0.0; fz[p] = exh[p]; }           /* :383-391 */
    FULL(dby);
    HL { /* :395-408 */
        size_t p = I3(i, j, k); double r = xm[I1(i)];
        fx[p] = eyh[p]; fy[p] = -exh[p] / r; fz[p] = 0.0; ss[p] = -(fx[p] / r);
    }
    FULL(dbz); SRCF(dbz);
#undef HL
#undef SRCF
#undef FULL
    const double zero = 0.0;   /* :414-436 */
    for (int k = 1; k <= kx; k++)
        for (int j = 1; j <= jx; j++)
            for (int i = 1; i <= ix - 1; i++)
                qx[I3(i, j, k)] = qav * dxm[I1(i)] * fmaxF(zero, fabs(vx[I3(i + 1, j, k)] - vx[I3(i, j, k)]) - R4_1EM4);
    for (int k = 1; k <= kx; k++)
        for (int j = 1; j <= jx - 1; j++)
            for (int i = 1; i <= ix; i++)
                qy[I3(i, j, k)] = qav * dym[I1(j)] * fmaxF(zero, fabs(vy[I3(i, j + 1, k)] - vy[I3(i, j, k)]) - R4_1EM4);
    for (int k = 1; k <= kx - 1; k++)
        for (int j = 1; j <= jx; j++)
            for (int i = 1; i <= ix; i++)
                qz[I3(i, j, k)] = qav * dzm[I1(k)] * fmaxF(zero, fabs(vz[I3(i, j, k + 1)] - vz[I3(i, j, k)]) - R4_1EM4);
#define ARTV(u, du) mlwartv3(u, du, dt, qx, dxi, dxim, ix, qy, dyi, dyim, jx, qz, dzi, dzim, kx)
    ARTV(ro, dro); ARTV(ee, dee); ARTV(rx, drx); ARTV(ry, dry); ARTV(rz, drz);
    ARTV(bx, dbx); ARTV(by, dby); ARTV(bz, dbz);
#undef ARTV
    for (int k = 2; k <= kx - 1; k++) /* :458-485 */
        for (int j = 2; j <= jx - 1; j++)
            for (int i = 2; i <= ix - 1; i++) {
                size_t p = I3(i, j, k);
                ro[p] = ro[p] + dro[p]; ee[p] = ee[p] + dee[p]; rx[p] = rx[p] + drx[p]; ry[p] = ry[p] + dry[p];
                rz[p] = rz[p] + drz[p]; bx[p] = bx[p] + dbx[p]; by[p] = by[p] + dby[p]; bz[p] = bz[p] + dbz[p];
                vx[p] = rx[p] / ro[p];
                vy[p] = ry[p] / ro[p];
                vz[p] = rz[p] / ro[p];
                double vv = sq(vx[p]) + sq(vy[p]) + sq(vz[p]);
                double bb = sq(bx[p]) + sq(by[p]) + sq(bz[p]);
                pr[p] = (gm - 1.0) * (ee[p] - 0.5 * ro[p] * vv - pi8i * bb);
            }
    double *tofree[] = {dxi, dxim, ux0, ux1, dyi, dyim, uy0, uy1, dzi, dzim, uz0, uz1, ex, ey, ez, ee, rx, ry, rz,
                        roh, eeh, rxh, ryh, rzh, bxh, byh, bzh, exh, eyh, ezh, prh, vxh, vyh, vzh,
                        dro, dee, drx, dry, drz, dbx, dby, dbz, fx, fy, fz, ss, qx, qy, qz};
    for (size_t t = 0; t < sizeof(tofree) / sizeof(tofree[0]); t++) free(tofree[t]);
}

/* hdmlw/mlw_m_s.f:2-589 (3D spherical (r, theta, phi) MHD, 8 fields): the
 * conserved variables and fluxes carry the metric factors (r^2 sin(theta),
 * r sin(theta), r, ...), the momentum equations the geometric sources
 * (mlwsrch/mlwsrcf); sin, cos, tan of y and ym from the C library, as the
 * reference's compiled intrinsics */
void or3d_mlw_m_s(double *ro, double *pr, double *vx, double *vy, double *vz,
                  double *bx, double *by, double *bz, double dt, double qav,
                  double gm, const double *x, const double *xm, const double *dx, const double *dxm, int ix,
                  const double *y, const double *ym, const double *dy, const double *dym, int jx,
                  const double *dz, const double *dzm, int kx) {
    const double pi = acos(-1.0);
    const double pi8i = 1.0 / pi / 8.0;
    const double pi4i = 1.0 / pi / 4.0;
    size_t n = (size_t)ix * jx * kx;
    double *dxi = dalloc(ix), *dxim = dalloc(ix), *ux0 = dalloc(ix), *ux1 = dalloc(ix);
    double *dyi = dalloc(jx), *dyim = dalloc(jx), *uy0 = dalloc(jx), *uy1 = dalloc(jx);
    double *dzi = dalloc(kx), *dzim = dalloc(kx), *uz0 = dalloc(kx), *uz1 = dalloc(kx);
    double *siny = dalloc(jx), *sinym = dalloc(jx);
    double *ex = dalloc(n), *ey = dalloc(n), *ez = dalloc(n);
    double *ee = dalloc(n), *rx = dalloc(n), *ry = dalloc(n), *rz = dalloc(n);
    double *sc[8], *sch[8], *dsc[8];
    for (int q = 0; q < 8; q++) { sc[q] = dalloc(n); sch[q] = dalloc(n); dsc[q] = dalloc(n); }
    double *roh = dalloc(n), *eeh = dalloc(n), *prh = dalloc(n), *vxh = dalloc(n), *vyh = dalloc(n), *vzh = dalloc(n);
    double *bxh = dalloc(n), *byh = dalloc(n), *bzh = dalloc(n), *exh = dalloc(n), *eyh = dalloc(n), *ezh = dalloc(n);
    double *fx = dalloc(n), *fy = dalloc(n), *fz = dalloc(n), *ss = dalloc(n);
    double *qx = dalloc(n), *qy = dalloc(n), *qz = dalloc(n);

    for (int i = 1; i <= ix; i++) { dxi[I1(i)] = 1.0 / dx[I1(i)]; dxim[I1(i)] = 1.0 / dxm[I1(i)]; }
    for (int i = 2; i <= ix - 1; i++) { ux1[I1(i)] = 0.5 * dxm[I1(i - 1)] / dx[I1(i)]; ux0[I1(i)] = 0.5 * dxm[I1(i)] / dx[I1(i)]; }
    for (int j = 1; j <= jx; j++) { dyi[I1(j)] = 1.0 / dy[I1(j)]; dyim[I1(j)] = 1.0 / dym[I1(j)]; }
    for (int j = 2; j <= jx - 1; j++) { uy1[I1(j)] = 0.5 * dym[I1(j - 1)] / dy[I1(j)]; uy0[I1(j)] = 0.5 * dym[I1(j)] / dy[I1(j)]; }
    for (int k = 1; k <= kx; k++) { dzi[I1(k)] = 1.0 / dz[I1(k)]; dzim[I1(k)] = 1.0 / dzm[I1(k)]; }
    for (int k = 2; k <= kx - 1; k++) { uz1[I1(k)] = 0.5 * dzm[I1(k - 1)] / dz[I1(k)]; uz0[I1(k)] = 0.5 * dzm[I1(k)] / dz[I1(k)]; }
    for (int j = 1; j <= jx; j++) { siny[I1(j)] = sin(y[I1(j)]); sinym[I1(j)] = sin(ym[I1(j)]); }   /* :109-112 */
    for (int q = 0; q < 8; q++)   /* :117-130 */
        for (size_t p = 0; p < n; p++) dsc[q][p] = 0.0;
    for (size_t p = 0; p < n; p++) {
        double vv = sq(vx[p]) + sq(vy[p]) + sq(vz[p]);
        double bb = sq(bx[p]) + sq(by[p]) + sq(bz[p]);
        ee[p] = pr[p] / (gm - 1.0) + 0.5 * ro[p] * vv + pi8i * bb;
        rx[p] = ro[p] * vx[p];
        ry[p] = ro[p] * vy[p];
        rz[p] = ro[p] * vz[p];
        ex[p] = -vy[p] * bz[p] + vz[p] * by[p];
        ey[p] = -vz[p] * bx[p] + vx[p] * bz[p];
        ez[p] = -vx[p] * by[p] + vy[p] * bx[p];
    }
#define CELLS for (int k = 1; k <= kx; k++) for (int j = 1; j <= jx; j++) for (int i = 1; i <= ix; i++)
    CELLS { /* :158-170 */
        size_t p = I3(i, j, k); double r = x[I1(i)], st = siny[I1(j)];
        sc[0][p] = ro[p] * sq(r) * st; sc[1][p] = ee[p] * sq(r) * st; sc[2][p] = rx[p] * sq(r) * st;
        sc[3][p] = ry[p] * r * st;      sc[4][p] = rz[p] * r;          sc[5][p] = bx[p] * sq(r) * st;
        sc[6][p] = by[p] * r * st;      sc[7][p] = bz[p] * r;
    }
#define HALF(q) mlwhalf3(sc[q], sch[q], dsc[q], dt, fx, dxi, dxim, ix, fy, dyi, dyim, jx, fz, dzi, dzim, kx)
#define SCALE3 fx[p] = fx[p] * sq(r) * st; fy[p] = fy[p] * r * st; fz[p] = fz[p] * r;
    CELLS { /* :172-189 */
        size_t p = I3(i, j, k); double r = x[I1(i)], st = siny[I1(j)];
        fx[p] = ro[p] * vx[p]; fy[p] = ro[p] * vy[p]; fz[p] = ro[p] * vz[p];
        SCALE3
    }
    HALF(0);
    CELLS { /* :191-210 */
        size_t p = I3(i, j, k); double r = x[I1(i)], st = siny[I1(j)];
        double vv = sq(vx[p]) + sq(vy[p]) + sq(vz[p]);
        double ep = pr[p] * gm / (gm - 1.0) + 0.5 * ro[p] * vv;
        double px = (bz[p] * ey[p] - by[p] * ez[p]) * pi4i;
        double py = (bx[p] * ez[p] - bz[p] * ex[p]) * pi4i;
        double pz = (by[p] * ex[p] - bx[p] * ey[p]) * pi4i;
        fx[p] = ep * vx[p] + px; fy[p] = ep * vy[p] + py; fz[p] = ep * vz[p] + pz;
        SCALE3
    }
    HALF(1);
    CELLS { /* :212-233 */
        size_t p = I3(i, j, k); double r = x[I1(i)], st = siny[I1(j)];
        fx[p] = ro[p] * sq(vx[p]) + pr[p] + pi8i * (sq(by[p]) + sq(bz[p]) - sq(bx[p]));
        fy[p] = ro[p] * vx[p] * vy[p] - pi4i * bx[p] * by[p];
        fz[p] = ro[p] * vx[p] * vz[p] - pi4i * bx[p] * bz[p];
        SCALE3
        ss[p] = (ro[p] * (sq(vy[p]) + sq(vz[p])) + 2.0 * pr[p] + pi4i * sq(bx[p])) * r * st;
    }
    HALF(2); mlwsrch3(sch[2], dsc[2], dt, ss, ix, jx, kx);
    CELLS { /* :236-258 */
        size_t p = I3(i, j, k); double r = x[I1(i)], st = siny[I1(j)];
        fx[p] = ro[p] * vy[p] * vx[p] - pi4i * by[p] * bx[p];
        fy[p] = ro[p] * sq(vy[p]) + pr[p] + pi8i * (sq(bx[p]) + sq(bz[p]) - sq(by[p]));
        fz[p] = ro[p] * vy[p] * vz[p] - pi4i * by[p] * bz[p];
        fx[p] = fx[p] * r * st;
        fy[p] = fy[p] * st;
        ss[p] = (ro[p] * vx[p] * vy[p] - pi4i * bx[p] * by[p]) * (-2.0) * st
              + (ro[p] * sq(vz[p]) + pr[p] + pi8i * (sq(bx[p]) + sq(by[p]) - sq(bz[p]))) * cos(y[I1(j)]);
    }
    HALF(3); mlwsrch3(sch[3], dsc[3], dt, ss, ix, jx, kx);
    CELLS { /* :261-282 */
        size_t p = I3(i, j, k); double r = x[I1(i)], st = siny[I1(j)];
        fx[p] = ro[p] * vz[p] * vx[p] - pi4i * bz[p] * bx[p];
        fy[p] = ro[p] * vz[p] * vy[p] - pi4i * bz[p] * by[p];
        fz[p] = ro[p] * sq(vz[p]) + pr[p] + pi8i * (sq(bx[p]) + sq(by[p]) - sq(bz[p]));
        fx[p] = fx[p] * r;
        fz[p] = fz[p] / st;
        ss[p] = (ro[p] * vx[p] * vz[p] - pi4i * bx[p] * bz[p]) * (-2.0)
              + (ro[p] * vy[p] * vz[p] - pi4i * by[p] * bz[p]) * (-2.0) / tan(y[I1(j)]);
    }
    HALF(4); mlwsrch3(sch[4], dsc[4], dt, ss, ix, jx, kx);
    CELLS { size_t p = I3(i, j, k); double r = x[I1(i)], st = siny[I1(j)];     /* :284-295 */
            fx[p] = 0.0; fy[p] = ez[p] * r * st; fz[p] = -ey[p] * r; }
    HALF(5);
    CELLS { size_t p = I3(i, j, k); double r = x[I1(i)], st = siny[I1(j)];     /* :297-308 */
            fx[p] = -ez[p] * r * st; fy[p] = 0.0; fz[p] = ex[p]; }
    HALF(6);
    CELLS { size_t p = I3(i, j, k); double r = x[I1(i)];                       /* :310-321 */
            fx[p] = ey[p] * r; fy[p] = -ex[p]; fz[p] = 0.0; }
    HALF(7);
#undef HALF
#undef CELLS
    for (int k = 1; k <= kx - 1; k++) /* :325-352 */
        for (int j = 1; j <= jx - 1; j++)
            for (int i = 1; i <= ix - 1; i++) {
                size_t p = I3(i, j, k); double r = xm[I1(i)], st = sinym[I1(j)];
                roh[p] = sch[0][p] / sq(r) / st;
                eeh[p] = sch[1][p] / sq(r) / st;
                vxh[p] = sch[2][p] / roh[p] / sq(r) / st;
                vyh[p] = sch[3][p] / roh[p] / r / st;
                vzh[p] = sch[4][p] / roh[p] / r;
                bxh[p] = sch[5][p] / sq(r) / st;
                byh[p] = sch[6][p] / r / st;
                bzh[p] = sch[7][p] / r;
                double vv = sq(vxh[p]) + sq(vyh[p]) + sq(vzh[p]);
                double bb = sq(bxh[p]) + sq(byh[p]) + sq(bzh[p]);
                prh[p] = (gm - 1.0) * (eeh[p] - 0.5 * roh[p] * vv - pi8i * bb);
                exh[p] = -vyh[p] * bzh[p] + vzh[p] * byh[p];
                eyh[p] = -vzh[p] * bxh[p] + vxh[p] * bzh[p];
                ezh[p] = -vxh[p] * byh[p] + vyh[p] * bxh[p];
            }
#define FULL(q) mlwfull3(dsc[q], dt, fx, dxi, ux0, ux1, ix, fy, dyi, uy0, uy1, jx, fz, dzi, uz0, uz1, kx)
#define SRCF(q) mlwsrcf3(dsc[q], dt, ss, ux0, ux1, ix, uy0, uy1, jx, uz0, uz1, kx)
#define HL for (int k = 1; k <= kx - 1; k++) for (int j = 1; j <= jx - 1; j++) for (int i = 1; i <= ix - 1; i++)
    HL { size_t p = I3(i, j, k); double r = xm[I1(i)], st = sinym[I1(j)];    /* :356-370 */
         fx[p] = roh[p] * vxh[p]; fy[p] = roh[p] * vyh[p]; fz[p] = roh[p] * vzh[p]; SCALE3 }
    FULL(0);
    HL { /* :373-391 */
        size_t p = I3(i, j, k); double r = xm[I1(i)], st = sinym[I1(j)];
        double vv = sq(vxh[p]) + sq(vyh[p]) + sq(vzh[p]);
        double ep = prh[p] * gm / (gm - 1.0) + 0.5 * roh[p] * vv;
        double px = (bzh[p] * eyh[p] - byh[p] * ezh[p]) * pi4i;
        double py = (bxh[p] * ezh[p] - bzh[p] * exh[p]) * pi4i;
        double pz = (byh[p] * exh[p] - bxh[p] * eyh[p]) * pi4i;
        fx[p] = ep * vxh[p] + px; fy[p] = ep * vyh[p] + py; fz[p] = ep * vzh[p] + pz;
        SCALE3
    }
    FULL(1);
    HL { /* :394-415 */
        size_t p = I3(i, j, k); double r = xm[I1(i)], st = sinym[I1(j)];
        fx[p] = roh[p] * sq(vxh[p]) + prh[p] + pi8i * (sq(byh[p]) + sq(bzh[p]) - sq(bxh[p]));
        fy[p] = roh[p] * vxh[p] * vyh[p] - pi4i * bxh[p] * byh[p];
        fz[p] = roh[p] * vxh[p] * vzh[p] - pi4i * bxh[p] * bzh[p];
        SCALE3
        ss[p] = (roh[p] * (sq(vyh[p]) + sq(vzh[p])) + 2.0 * prh[p] + pi4i * sq(bxh[p])) * r * st;
    }
    FULL(2); SRCF(2);
    HL { /* :418-440 */
        size_t p = I3(i, j, k); double r = xm[I1(i)], st = sinym[I1(j)];
        fx[p] = roh[p] * vyh[p] * vxh[p] - pi4i * byh[p] * bxh[p];
        fy[p] = roh[p] * sq(vyh[p]) + prh[p] + pi8i * (sq(bxh[p]) + sq(bzh[p]) - sq(byh[p]));
        fz[p] = roh[p] * vyh[p] * vzh[p] - pi4i * byh[p] * bzh[p];
        fx[p] = fx[p] * r * st;
        fy[p] = fy[p] * st;
        ss[p] = (roh[p] * vxh[p] * vyh[p] - pi4i * bxh[p] * byh[p]) * (-2.0) * st
              + (roh[p] * sq(vzh[p]) + prh[p] + pi8i * (sq(bxh[p]) + sq(byh[p]) - sq(bzh[p]))) * cos(ym[I1(j)]);
    }
    FULL(3); SRCF(3);
    HL { /* :443-464 */
        size_t p = I3(i, j, k); double r = xm[I1(i)], st = sinym[I1(j)];
        fx[p] = roh[p] * vzh[p] * vxh[p] - pi4i * bzh[p] * bxh[p];
        fy[p] = roh[p] * vzh[p] * vyh[p] - pi4i * bzh[p] * byh[p];
        fz[p] = roh[p] * sq(vzh[p]) + prh[p] + pi8i * (sq(bxh[p]) + sq(byh[p]) - sq(bzh[p]));
        fx[p] = fx[p] * r;
        fz[p] = fz[p] / st;
        ss[p] = (roh[p] * vxh[p] * vzh[p] - pi4i * bxh[p] * bzh[p]) * (-2.0)
              + (roh[p] * vyh[p] * vzh[p] - pi4i * byh[p] * bzh[p]) * (-2.0) / tan(ym[I1(j)]);
    }
    FULL(4); SRCF(4);
    HL { size_t p = I3(i, j, k); double r = xm[I1(i)], st = sinym[I1(j)];     /* :466-477 */
         fx[p] = 0.0; fy[p] = ezh[p] * r * st; fz[p] = -eyh[p] * r; }
    FULL(5);
    HL { size_t p = I3(i, j, k); double r = xm[I1(i)], st = sinym[I1(j)];     /* :479-490 */
         fx[p] = -ezh[p] * r * st; fy[p] = 0.0; fz[p] = exh[p]; }
    FULL(6);
    HL { size_t p = I3(i, j, k); double r = xm[I1(i)];                        /* :492-503 */
         fx[p] = eyh[p] * r; fy[p] = -exh[p]; fz[p] = 0.0; }
    FULL(7);
#undef HL
#undef SRCF
#undef FULL
#undef SCALE3
    const double zero = 0.0;   /* :509-531 */
    for (int k = 1; k <= kx; k++)
        for (int j = 1; j <= jx; j++)
            for (int i = 1; i <= ix - 1; i++)
                qx[I3(i, j, k)] = qav * dxm[I1(i)] * fmaxF(zero, fabs(vx[I3(i + 1, j, k)] - vx[I3(i, j, k)]) - R4_1EM4);
    for (int k = 1; k <= kx; k++)
        for (int j = 1; j <= jx - 1; j++)
            for (int i = 1; i <= ix; i++)
                qy[I3(i, j, k)] = qav * dym[I1(j)] * fmaxF(zero, fabs(vy[I3(i, j + 1, k)] - vy[I3(i, j, k)]) - R4_1EM4);
    for (int k = 1; k <= kx - 1; k++)
        for (int j = 1; j <= jx; j++)
            for (int i = 1; i <= ix; i++)
                qz[I3(i, j, k)] = qav * dzm[I1(k)] * fmaxF(zero, fabs(vz[I3(i, j, k + 1)] - vz[I3(i, j, k)]) - R4_1EM4);
    for (int q = 0; q < 8; q++) /* :534-549 */
        mlwartv3(sc[q], dsc[q], dt, qx, dxi, dxim, ix, qy, dyi, dyim, jx, qz, dzi, dzim, kx);
    for (int k = 2; k <= kx - 1; k++) /* :553-586 */
        for (int j = 2; j <= jx - 1; j++)
            for (int i = 2; i <= ix - 1; i++) {
                size_t p = I3(i, j, k); double r = x[I1(i)], st = siny[I1(j)];
                double u[8];
                for (int q = 0; q < 8; q++) u[q] = sc[q][p] + dsc[q][p];
                ro[p] = u[0] / sq(r) / st;
                double e = u[1] / sq(r) / st;
                vx[p] = u[2] / ro[p] / sq(r) / st;
                vy[p] = u[3] / ro[p] / r / st;
                vz[p] = u[4] / ro[p] / r;
                bx[p] = u[5] / sq(r) / st;
                by[p] = u[6] / r / st;
                bz[p] = u[7] / r;
                double vv = sq(vx[p]) + sq(vy[p]) + sq(vz[p]);
                double bb = sq(bx[p]) + sq(by[p]) + sq(bz[p]);
                pr[p] = (gm - 1.0) * (e - 0.5 * ro[p] * vv - pi8i * bb);
            }
    double *tofree[] = {dxi, dxim, ux0, ux1, dyi, dyim, uy0, uy1, dzi, dzim, uz0, uz1, siny, sinym, ex, ey, ez, ee, rx,
                        ry, rz, roh, eeh, prh, vxh, vyh, vzh, bxh, byh, bzh, exh, eyh, ezh, fx, fy, fz, ss, qx, qy, qz};
    for (size_t t = 0; t < sizeof(tofree) / sizeof(tofree[0]); t++) free(tofree[t]);
    for (int q = 0; q < 8; q++) { free(sc[q]); free(sch[q]); free(dsc[q]); }
}

/* common/cfl_m.f:39-80 (all cells; REAL*4 pi) */
void or3d_cfl_m(double *dt, double safety, double dtmin, int *merr, double gm,
                const double *ro, const double *pr, const double *vx,
                const double *vy, const double *vz, const double *bx,
                const double *by, const double *bz, const double *dx, int ix,
                const double *dy, int jx, const double *dz, int kx,
                int *imin, int *jmin, int *kmin) {
    const double pai = (double)acosf(-1.0f);
    const double onep4 = 0.25 / pai;
    size_t n = (size_t)ix * jx * kx;
    double *dtq = dalloc(n);
    double d = R4_1E20;
    int im = 0, jm = 0, km = 0;
    for (int k = 1; k <= kx; k++)
        for (int j = 1; j <= jx; j++)
            for (int i = 1; i <= ix; i++) {
                size_t p = I3(i, j, k);
                double b2 = sq(bx[p]) + sq(by[p]) + sq(bz[p]);
                double ca2 = b2 * onep4 / ro[p];
                double cs2 = gm * pr[p] / ro[p];
                double v2 = sq(vx[p]) + sq(vy[p]) + sq(vz[p]);
                double dtcfl = fminF(fminF(dx[I1(i)], dy[I1(j)]), dz[I1(k)]) / sqrt(v2 + cs2 + ca2);
                dtq[p] = safety * dtcfl;
            }
    for (int k = 1; k <= kx; k++)
        for (int j = 1; j <= jx; j++)
            for (int i = 1; i <= ix; i++)
                if (dtq[I3(i, j, k)] < d) { im = i; jm = j; km = k; d = dtq[I3(i, j, k)]; }
    *merr = (d < dtmin) ? 9001 : 0;
    *dt = d;
    if (imin) *imin = im;
    if (jmin) *jmin = jm;
    if (kmin) *kmin = km;
    free(dtq);
}

/* common/cfl_m_s.f:1-50: cfl_m with the widths dx, x dy, x sin(y) dz */
void or3d_cfl_m_s(double *dt, double safety, double dtmin, int *merr, double gm,
                  const double *ro, const double *pr, const double *vx,
                  const double *vy, const double *vz, const double *bx,
                  const double *by, const double *bz, const double *x, const double *dx, int ix,
                  const double *y, const double *dy, int jx, const double *dz, int kx,
                  int *imin, int *jmin, int *kmin) {
    const double pai = (double)acosf(-1.0f);
    const double onep4 = 0.25 / pai;
    double d = R4_1E20;
    int im = 0, jm = 0, km = 0;
    for (int k = 1; k <= kx; k++)   /* :19-40, the two loops fused in the same order */
        for (int j = 1; j <= jx; j++)
            for (int i = 1; i <= ix; i++) {
                size_t p = I3(i, j, k);
                double b2 = sq(bx[p]) + sq(by[p]) + sq(bz[p]);
                double ca2 = b2 * onep4 / ro[p];
                double cs2 = gm * pr[p] / ro[p];
                double v2 = sq(vx[p]) + sq(vy[p]) + sq(vz[p]);
                double dtcfl = fminF(fminF(dx[I1(i)], x[I1(i)] * dy[I1(j)]), x[I1(i)] * sin(y[I1(j)]) * dz[I1(k)])
                               / sqrt(v2 + cs2 + ca2);
                double q = safety * dtcfl;
                if (q < d) { im = i; jm = j; km = k; d = q; }
            }
    *merr = (d < dtmin) ? 9001 : 0;
    *dt = d;
    if (imin) *imin = im;
    if (jmin) *jmin = jm;
    if (kmin) *kmin = km;
}

/* bc/bd*{x,y,z}.f (3D): the 1D index map along `dir` for every line */
void or3d_bd(int dir, int kind, int mbnd, int margin, double *qq, double q0,
             const double *dm, int ix, int jx, int kx) {
    int n = dir == 0 ? ix : (dir == 1 ? jx : kx);
    size_t stride = dir == 0 ? 1 : (dir == 1 ? (size_t)ix : (size_t)ix * jx);
    double *line = dalloc(n);
    for (int k = 1; k <= (dir == 2 ? 1 : kx); k++)
        for (int j = 1; j <= (dir == 1 ? 1 : jx); j++)
            for (int i = 1; i <= (dir == 0 ? 1 : ix); i++) {
                size_t base = I3(i, j, k);
                for (int a = 0; a < n; a++) line[a] = qq[base + (size_t)a * stride];
                if (kind == 7) or1d_bdperx(mbnd, margin, line, n);
                else or1d_bd(kind, mbnd, margin, line, q0, dm, n);
                for (int a = 0; a < n; a++) qq[base + (size_t)a * stride] = line[a];
            }
    free(line);
}

/* ---- the helper routines themselves, with the reference argument lists */
void or3d_mlwhalf(const double *u, double *un, double *du, double dt, const double *fx, const double *dxi,
                  const double *dxim, int ix, const double *fy, const double *dyi, const double *dyim, int jx,
                  const double *fz, const double *dzi, const double *dzim, int kx) {
    mlwhalf3(u, un, du, dt, fx, dxi, dxim, ix, fy, dyi, dyim, jx, fz, dzi, dzim, kx);
}
void or3d_mlwfull(double *du, double dt, const double *fx, const double *dxi, const double *ux0, const double *ux1,
                  int ix, const double *fy, const double *dyi, const double *uy0, const double *uy1, int jx,
                  const double *fz, const double *dzi, const double *uz0, const double *uz1, int kx) {
    mlwfull3(du, dt, fx, dxi, ux0, ux1, ix, fy, dyi, uy0, uy1, jx, fz, dzi, uz0, uz1, kx);
}
void or3d_mlwartv(const double *u, double *du, double dt, const double *qx, const double *dxi, const double *dxim,
                  int ix, const double *qy, const double *dyi, const double *dyim, int jx, const double *qz,
                  const double *dzi, const double *dzim, int kx) {
    mlwartv3(u, du, dt, qx, dxi, dxim, ix, qy, dyi, dyim, jx, qz, dzi, dzim, kx);
}
void or3d_mlwsrch(double *un, double *du, double dt, const double *s, int ix, int jx, int kx) {
    mlwsrch3(un, du, dt, s, ix, jx, kx);
}
void or3d_mlwsrcf(double *du, double dt, const double *s, const double *ux0, const double *ux1, int ix,
                  const double *uy0, const double *uy1, int jx, const double *uz0, const double *uz1, int kx) {
    mlwsrcf3(du, dt, s, ux0, ux1, ix, uy0, uy1, jx, uz0, uz1, kx);
}
