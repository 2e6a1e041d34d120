/*
 * cans_oracle2d.c -- C restatement of the cans2d hot path (test
 * infrastructure only; see cans_oracle.h).  Citations are to
 * cans_beta_vary/cans2d/.
 */
#include "cans_oracle.h"
#include "cans_oracle_util.h"

/* hdmlw/mlwhalf.f:36-52 (corner-staggered half step) */
static void mlwhalf2(const double *u, double *un, double *du, double dt,
                     const double *fx, const double *dxi, const double *dxim, int ix,
                     const double *fy, const double *dyi, const double *dyim, int jx) {
    for (int j = 2; j <= jx - 1; j++)
        for (int i = 2; i <= ix - 1; i++)
            du[I2(i, j)] = du[I2(i, j)] - 0.5 * dt * (0.5 * dxi[I1(i)] * (fx[I2(i + 1, j)] - fx[I2(i - 1, j)]))
                                        - 0.5 * dt * (0.5 * dyi[I1(j)] * (fy[I2(i, j + 1)] - fy[I2(i, j - 1)]));
    for (int j = 1; j <= jx - 1; j++)
        for (int i = 1; i <= ix - 1; i++) {
            double uh = 0.25 * (u[I2(i + 1, j)] + u[I2(i, j)] + u[I2(i + 1, j + 1)] + u[I2(i, j + 1)]);
            double dfdx = dxim[I1(i)] * (fx[I2(i + 1, j)] - fx[I2(i, j)] + fx[I2(i + 1, j + 1)] - fx[I2(i, j + 1)]) / 2.0;
            double dfdy = dyim[I1(j)] * (fy[I2(i, j + 1)] - fy[I2(i, j)] + fy[I2(i + 1, j + 1)] - fy[I2(i + 1, j)]) / 2.0;
            un[I2(i, j)] = uh - dt * dfdx - dt * dfdy;
        }
}

/* hdmlw/mlwfull.f:31-41 */
static void mlwfull2(double *du, double dt, const double *fx, const double *dxi,
                     const double *ux0, const double *ux1, int ix,
                     const double *fy, const double *dyi, const double *uy0,
                     const double *uy1, int jx) {
    for (int j = 2; j <= jx - 1; j++)
        for (int i = 2; i <= ix - 1; i++) {
            double dfdx = dxi[I1(i)] * (uy0[I1(j)] * (fx[I2(i, j - 1)] - fx[I2(i - 1, j - 1)])
                                        + uy1[I1(j)] * (fx[I2(i, j)] - fx[I2(i - 1, j)]));
            double dfdy = dyi[I1(j)] * (ux0[I1(i)] * (fy[I2(i - 1, j)] - fy[I2(i - 1, j - 1)])
                                        + ux1[I1(i)] * (fy[I2(i, j)] - fy[I2(i, j - 1)]));
            du[I2(i, j)] = du[I2(i, j)] - 0.5 * dt * dfdx - 0.5 * dt * dfdy;
        }
}

/* hdmlw/mlwartv.f:33-41 */
static void mlwartv2(const double *u, double *du, double dt, const double *qx,
                     const double *dxi, const double *dxim, int ix,
                     const double *qy, const double *dyi, const double *dyim, int jx) {
    for (int j = 2; j <= jx - 1; j++)
        for (int i = 2; i <= ix - 1; i++)
            du[I2(i, j)] = du[I2(i, j)]
                + dt * (dxi[I1(i)] * (qx[I2(i, j)] * dxim[I1(i)] * (u[I2(i + 1, j)] - u[I2(i, j)])
                                      - qx[I2(i - 1, j)] * dxim[I1(i - 1)] * (u[I2(i, j)] - u[I2(i - 1, j)])))
                + dt * (dyi[I1(j)] * (qy[I2(i, j)] * dyim[I1(j)] * (u[I2(i, j + 1)] - u[I2(i, j)])
                                      - qy[I2(i, j - 1)] * dyim[I1(j - 1)] * (u[I2(i, j)] - u[I2(i, j - 1)])));
}

/* hdmlw/mlwsrch.f:32-47 */
static void mlwsrch2(double *un, double *du, double dt, const double *s, int ix, int jx) {
    for (int j = 2; j <= jx - 1; j++)
        for (int i = 2; i <= ix - 1; i++)
            du[I2(i, j)] = du[I2(i, j)] + 0.5 * dt * s[I2(i, j)];
    for (int j = 1; j <= jx - 1; j++)
        for (int i = 1; i <= ix - 1; i++) {
            double sh = (s[I2(i + 1, j)] + s[I2(i, j)] + s[I2(i + 1, j + 1)] + s[I2(i, j + 1)]) / 4.0;
            un[I2(i, j)] = un[I2(i, j)] + dt * sh;
        }
}

/* hdmlw/mlwsrcf.f:33-40 */
static void mlwsrcf2(double *du, double dt, const double *s, const double *ux0,
                     const double *ux1, int ix, const double *uy0,
                     const double *uy1, int jx) {
    for (int j = 2; j <= jx - 1; j++)
        for (int i = 2; i <= ix - 1; i++) {
            double sh = uy0[I1(j)] * (ux0[I1(i)] * s[I2(i - 1, j - 1)] + ux1[I1(i)] * s[I2(i, j - 1)])
                      + uy1[I1(j)] * (ux0[I1(i)] * s[I2(i - 1, j)] + ux1[I1(i)] * s[I2(i, j)]);
            du[I2(i, j)] = du[I2(i, j)] + 0.5 * dt * sh;
        }
}

/* hdmlw/mlw_m.f:2-324 (2D MHD with vector potential Az) */
void or2d_mlw_m(double *ro, double *pr, double *vx, double *vy, double *bx,
                double *by, double *az, double dt, double qav, double gm,
                const double *dx, const double *dxm, int ix,
                const double *dy, const double *dym, int jx) {
    const double pi = acos(-1.0);
    const double pi8i = 1.0 / pi / 8.0;
    const double pi4i = 1.0 / pi / 4.0;
    size_t n = (size_t)ix * jx;
    double *dxi = dalloc(ix), *dxim = dalloc(ix), *ux0 = dalloc(ix), *ux1 = dalloc(ix);
    double *dyi = dalloc(jx), *dyim = dalloc(jx), *uy0 = dalloc(jx), *uy1 = dalloc(jx);
    double *ee = dalloc(n), *rx = dalloc(n), *ry = dalloc(n), *ez = dalloc(n);
    double *roh = dalloc(n), *eeh = dalloc(n), *rxh = dalloc(n), *ryh = dalloc(n), *bxh = dalloc(n), *byh = dalloc(n);
    double *prh = dalloc(n), *vxh = dalloc(n), *vyh = dalloc(n), *ezh = dalloc(n), *azh = dalloc(n);
    double *dro = dalloc(n), *dee = dalloc(n), *drx = dalloc(n), *dry = dalloc(n), *dbx = dalloc(n), *dby = dalloc(n), *daz = dalloc(n);
    double *fx = dalloc(n), *fy = dalloc(n), *qx = dalloc(n), *qy = dalloc(n), *ss = dalloc(n);

    for (int i = 1; i <= ix; i++) { dxi[I1(i)] = 1.0 / dx[I1(i)]; dxim[I1(i)] = 1.0 / dxm[I1(i)]; } /* :68-71 */
    for (int i = 2; i <= ix - 1; i++) {                                                             /* :72-75 */
        ux1[I1(i)] = 0.5 * dxm[I1(i - 1)] / dx[I1(i)];
        ux0[I1(i)] = 0.5 * dxm[I1(i)] / dx[I1(i)];
    }
    for (int j = 1; j <= jx; j++) { dyi[I1(j)] = 1.0 / dy[I1(j)]; dyim[I1(j)] = 1.0 / dym[I1(j)]; }
    for (int j = 2; j <= jx - 1; j++) {
        uy1[I1(j)] = 0.5 * dym[I1(j - 1)] / dy[I1(j)];
        uy0[I1(j)] = 0.5 * dym[I1(j)] / dy[I1(j)];
    }
    for (size_t p = 0; p < n; p++) { dro[p] = 0.0; dee[p] = 0.0; drx[p] = 0.0; dry[p] = 0.0; dbx[p] = 0.0; dby[p] = 0.0; daz[p] = 0.0; }
    for (int j = 1; j <= jx; j++)
        for (int i = 1; i <= ix; i++) { /* :101-110 */
            size_t p = I2(i, j);
            double vv = sq(vx[p]) + sq(vy[p]);
            double bb = sq(bx[p]) + sq(by[p]);
            ee[p] = pr[p] / (gm - 1.0) + 0.5 * ro[p] * vv + pi8i * bb;
            rx[p] = ro[p] * vx[p];
            ry[p] = ro[p] * vy[p];
        }
    for (size_t p = 0; p < n; p++) ez[p] = -vx[p] * by[p] + vy[p] * bx[p];

    /* half step, :121-184 */
    for (size_t p = 0; p < n; p++) { fx[p] = ro[p] * vx[p]; fy[p] = ro[p] * vy[p]; }
    mlwhalf2(ro, roh, dro, dt, fx, dxi, dxim, ix, fy, dyi, dyim, jx);
    for (size_t p = 0; p < n; p++) {
        double vv = sq(vx[p]) + sq(vy[p]);
        double ep = pr[p] * gm / (gm - 1.0) + 0.5 * ro[p] * vv;
        fx[p] = ep * vx[p] - by[p] * ez[p] * pi4i;
        fy[p] = ep * vy[p] + bx[p] * ez[p] * pi4i;
    }
    mlwhalf2(ee, eeh, dee, dt, fx, dxi, dxim, ix, fy, dyi, dyim, jx);
    for (size_t p = 0; p < n; p++) {
        fx[p] = ro[p] * sq(vx[p]) + pr[p] + pi8i * (sq(by[p]) - sq(bx[p]));
        fy[p] = ro[p] * vx[p] * vy[p] - pi4i * bx[p] * by[p];
    }
    mlwhalf2(rx, rxh, drx, dt, fx, dxi, dxim, ix, fy, dyi, dyim, jx);
    for (size_t p = 0; p < n; p++) {
        fx[p] = ro[p] * vy[p] * vx[p] - pi4i * by[p] * bx[p];
        fy[p] = ro[p] * sq(vy[p]) + pr[p] + pi8i * (sq(bx[p]) - sq(by[p]));
    }
    mlwhalf2(ry, ryh, dry, dt, fx, dxi, dxim, ix, fy, dyi, dyim, jx);
    for (size_t p = 0; p < n; p++) { fx[p] = 0.0; fy[p] = ez[p]; }
    mlwhalf2(bx, bxh, dbx, dt, fx, dxi, dxim, ix, fy, dyi, dyim, jx);
    for (size_t p = 0; p < n; p++) { fx[p] = -ez[p]; fy[p] = 0.0; }
    mlwhalf2(by, byh, dby, dt, fx, dxi, dxim, ix, fy, dyi, dyim, jx);
    for (size_t p = 0; p < n; p++) ss[p] = -ez[p];
    mlwsrch2(azh, daz, dt, ss, ix, jx);

    /* half-step primitives, :189-202 */
    for (int j = 1; j <= jx - 1; j++)
        for (int i = 1; i <= ix - 1; i++) {
            size_t p = I2(i, j);
            vxh[p] = rxh[p] / roh[p];
            vyh[p] = ryh[p] / roh[p];
            double vv = sq(vxh[p]) + sq(vyh[p]);
            double bb = sq(bxh[p]) + sq(byh[p]);
            prh[p] = (gm - 1.0) * (eeh[p] - 0.5 * roh[p] * vv - pi8i * bb);
        }
    for (size_t p = 0; p < n; p++) ezh[p] = -vxh[p] * byh[p] + vyh[p] * bxh[p];

    /* full step, :207-270 */
#define HLOOP for (int j = 1; j <= jx - 1; j++) for (int i = 1; i <= ix - 1; i++)
    HLOOP { size_t p = I2(i, j); fx[p] = roh[p] * vxh[p]; fy[p] = roh[p] * vyh[p]; }
    mlwfull2(dro, dt, fx, dxi, ux0, ux1, ix, fy, dyi, uy0, uy1, jx);
    HLOOP {
        size_t p = I2(i, j);
        double vv = sq(vxh[p]) + sq(vyh[p]);
        double ep = prh[p] * gm / (gm - 1.0) + 0.5 * roh[p] * vv;
        fx[p] = ep * vxh[p] - byh[p] * ezh[p] * pi4i;
        fy[p] = ep * vyh[p] + bxh[p] * ezh[p] * pi4i;
    }
    mlwfull2(dee, dt, fx, dxi, ux0, ux1, ix, fy, dyi, uy0, uy1, jx);
    HLOOP {
        size_t p = I2(i, j);
        fx[p] = roh[p] * sq(vxh[p]) + prh[p] + pi8i * (sq(byh[p]) - sq(bxh[p]));
        fy[p] = roh[p] * vxh[p] * vyh[p] - pi4i * bxh[p] * byh[p];
    }
    mlwfull2(drx, dt, fx, dxi, ux0, ux1, ix, fy, dyi, uy0, uy1, jx);
    HLOOP {
        size_t p = I2(i, j);
        fx[p] = roh[p] * vyh[p] * vxh[p] - pi4i * byh[p] * bxh[p];
        fy[p] = roh[p] * sq(vyh[p]) + prh[p] + pi8i * (sq(bxh[p]) - sq(byh[p]));
    }
    mlwfull2(dry, dt, fx, dxi, ux0, ux1, ix, fy, dyi, uy0, uy1, jx);
    HLOOP { size_t p = I2(i, j); fx[p] = 0.0; fy[p] = ezh[p]; }
    mlwfull2(dbx, dt, fx, dxi, ux0, ux1, ix, fy, dyi, uy0, uy1, jx);
    HLOOP { size_t p = I2(i, j); fx[p] = -ezh[p]; fy[p] = 0.0; }
    mlwfull2(dby, dt, fx, dxi, ux0, ux1, ix, fy, dyi, uy0, uy1, jx);
#undef HLOOP
    for (size_t p = 0; p < n; p++) ss[p] = -ezh[p];
    mlwsrcf2(daz, dt, ss, ux0, ux1, ix, uy0, uy1, jx);

    /* artificial viscosity, :276-295 */
    const double zero = 0.0;
    for (int j = 1; j <= jx - 1; j++)
        for (int i = 1; i <= ix - 1; i++)
            qx[I2(i, j)] = qav * dxm[I1(i)] * fmaxF(zero, fabs(vx[I2(i + 1, j)] - vx[I2(i, j)]) - R4_1EM4);
    for (int j = 1; j <= jx - 1; j++)
        for (int i = 1; i <= ix; i++)
            qy[I2(i, j)] = qav * dym[I1(j)] * fmaxF(zero, fabs(vy[I2(i, j + 1)] - vy[I2(i, j)]) - R4_1EM4);
    mlwartv2(ro, dro, dt, qx, dxi, dxim, ix, qy, dyi, dyim, jx);
    mlwartv2(ee, dee, dt, qx, dxi, dxim, ix, qy, dyi, dyim, jx);
    mlwartv2(rx, drx, dt, qx, dxi, dxim, ix, qy, dyi, dyim, jx);
    mlwartv2(ry, dry, dt, qx, dxi, dxim, ix, qy, dyi, dyim, jx);
    mlwartv2(bx, dbx, dt, qx, dxi, dxim, ix, qy, dyi, dyim, jx);
    mlwartv2(by, dby, dt, qx, dxi, dxim, ix, qy, dyi, dyim, jx);

    /* update and primitive recovery, :299-321 */
    for (int j = 2; j <= jx - 1; j++)
        for (int i = 2; i <= ix - 1; i++) {
            size_t p = I2(i, j);
            ro[p] = ro[p] + dro[p];
            ee[p] = ee[p] + dee[p];
            rx[p] = rx[p] + drx[p];
            ry[p] = ry[p] + dry[p];
            bx[p] = bx[p] + dbx[p];
            by[p] = by[p] + dby[p];
            az[p] = az[p] + daz[p];
        }
    for (int j = 2; j <= jx - 1; j++)
        for (int i = 2; i <= ix - 1; i++) {
            size_t p = I2(i, j);
            vx[p] = rx[p] / ro[p];
            vy[p] = ry[p] / ro[p];
            double vv = sq(vx[p]) + sq(vy[p]);
            double bb = sq(bx[p]) + sq(by[p]);
            pr[p] = (gm - 1.0) * (ee[p] - 0.5 * ro[p] * vv - pi8i * bb);
        }
    double *tofree[] = {dxi, dxim, ux0, ux1, dyi, dyim, uy0, uy1, ee, rx, ry, ez, roh, eeh, rxh, ryh, bxh, byh,
                        prh, vxh, vyh, ezh, azh, dro, dee, drx, dry, dbx, dby, daz, fx, fy, qx, qy, ss};
    for (size_t t = 0; t < sizeof(tofree) / sizeof(tofree[0]); t++) free(tofree[t]);
}

/* common/cfl_m.f:38-71 (all cells, ghosts included) */
void or2d_cfl_m(double *dt, double safety, double dtmin, int *merr, double gm,
                const double *ro, const double *pr, const double *vx,
                const double *vy, const double *bx, const double *by,
                const double *dx, int ix, const double *dy, int jx,
                int *imin, int *jmin) {
    const double pi = acos(-1.0);
    const double pi4i = 1.0 / pi / 4.0;
    size_t n = (size_t)ix * jx;
    double *dtq = dalloc(n);
    double d = R4_1E20;
    int im = 0, jm = 0;
    for (int j = 1; j <= jx; j++)
        for (int i = 1; i <= ix; i++) {
            size_t p = I2(i, j);
            double b2 = bx[p] * bx[p] + by[p] * by[p];
            double ca2 = b2 * pi4i / ro[p];
            double cs2 = gm * pr[p] / ro[p];
            double v2 = vx[p] * vx[p] + vy[p] * vy[p];
            double dtcfl = fminF(dx[I1(i)], dy[I1(j)]) / sqrt(v2 + cs2 + ca2);
            dtq[p] = safety * dtcfl;
        }
    for (int j = 1; j <= jx; j++)
        for (int i = 1; i <= ix; i++)
            if (dtq[I2(i, j)] < d) { im = i; jm = j; d = dtq[I2(i, j)]; }
    *merr = (d < dtmin) ? 9001 : 0;
    *dt = d;
    if (imin) *imin = im;
    if (jmin) *jmin = jm;
    free(dtq);
}

/* bc/bd*{x,y}.f (2D): the 1D index maps applied along x for every j, or
 * along y for every i.  kind 7 = bdper{x,y} with (mbnd, margin) read as
 * (margin0, margin1). */
void or2d_bd(int dir, int kind, int mbnd, int margin, double *qq, double q0,
             const double *dm, int ix, int jx) {
    int n = dir == 0 ? ix : jx;
    int lines = dir == 0 ? jx : ix;
    double *line = dalloc(n);
    for (int l = 1; l <= lines; l++) {
        for (int a = 1; a <= n; a++) line[I1(a)] = dir == 0 ? qq[I2(a, l)] : qq[I2(l, a)];
        if (kind == 7) or1d_bdperx(mbnd, margin, line, n);
        else or1d_bd(kind, mbnd, margin, line, q0, dm, n);
        for (int a = 1; a <= n; a++) {
            if (dir == 0) qq[I2(a, l)] = line[I1(a)];
            else qq[I2(l, a)] = line[I1(a)];
        }
    }
    free(line);
}

/* common/bbtoaa.f:28-49 (vector potential from B, trapezoid rule) */
void or2d_bbtoaa(double *az, const double *bx, const double *by,
                 const double *dxm, const double *dym, int margin, int ix,
                 int jx) {
    az[I2(margin + 1, margin + 1)] = 0.0;
    int i = margin + 1;
    for (int j = margin + 2; j <= jx; j++)
        az[I2(i, j)] = az[I2(i, j - 1)] + 0.5 * (bx[I2(i, j - 1)] + bx[I2(i, j)]) * dym[I1(j - 1)];
    for (int j = margin; j >= 1; j--)
        az[I2(i, j)] = az[I2(i, j + 1)] - 0.5 * (bx[I2(i, j)] + bx[I2(i, j + 1)]) * dym[I1(j)];
    for (int j = 1; j <= jx; j++) {
        for (int ii = margin + 2; ii <= ix; ii++)
            az[I2(ii, j)] = az[I2(ii - 1, j)] - 0.5 * (by[I2(ii - 1, j)] + by[I2(ii, j)]) * dxm[I1(ii - 1)];
        for (int ii = margin; ii >= 1; ii--)
            az[I2(ii, j)] = az[I2(ii + 1, j)] + 0.5 * (by[I2(ii, j)] + by[I2(ii + 1, j)]) * dxm[I1(ii)];
    }
}

/* hdmlw/mlw_m3.f:2-377: as mlw_m with the third components vz, bz (2D
 * fields in 3D space); Az evolves by -Ez as in mlw_m */
void or2d_mlw_m3(double *ro, double *pr, double *vx, double *vy, double *vz,
                 double *bx, double *by, double *bz, double *az, double dt,
                 double qav, double gm, const double *dx, const double *dxm,
                 int ix, const double *dy, const double *dym, int jx) {
    const double pi = acos(-1.0);
    const double pi8i = 1.0 / pi / 8.0; /* :65 */
    const double pi4i = 1.0 / pi / 4.0;
    size_t n = (size_t)ix * jx;
    double *dxi = dalloc(ix), *dxim = dalloc(ix), *ux0 = dalloc(ix), *ux1 = dalloc(ix);
    double *dyi = dalloc(jx), *dyim = dalloc(jx), *uy0 = dalloc(jx), *uy1 = dalloc(jx);
    double *ee = dalloc(n), *rx = dalloc(n), *ry = dalloc(n), *rz = dalloc(n);
    double *ex = dalloc(n), *ey = dalloc(n), *ez = dalloc(n);
    double *roh = dalloc(n), *eeh = dalloc(n), *rxh = dalloc(n), *ryh = dalloc(n), *rzh = dalloc(n);
    double *bxh = dalloc(n), *byh = dalloc(n), *bzh = dalloc(n), *azh = dalloc(n);
    double *prh = dalloc(n), *vxh = dalloc(n), *vyh = dalloc(n), *vzh = dalloc(n);
    double *exh = dalloc(n), *eyh = dalloc(n), *ezh = dalloc(n);
    double *dro = dalloc(n), *dee = dalloc(n), *drx = dalloc(n), *dry = dalloc(n), *drz = dalloc(n);
    double *dbx = dalloc(n), *dby = dalloc(n), *dbz = dalloc(n), *daz = dalloc(n);
    double *fx = dalloc(n), *fy = dalloc(n), *qx = dalloc(n), *qy = dalloc(n), *ss = dalloc(n);

    for (int i = 1; i <= ix; i++) { dxi[I1(i)] = 1.0 / dx[I1(i)]; dxim[I1(i)] = 1.0 / dxm[I1(i)]; }
    for (int i = 2; i <= ix - 1; i++) {
        ux1[I1(i)] = 0.5 * dxm[I1(i - 1)] / dx[I1(i)];
        ux0[I1(i)] = 0.5 * dxm[I1(i)] / dx[I1(i)];
    }
    for (int j = 1; j <= jx; j++) { dyi[I1(j)] = 1.0 / dy[I1(j)]; dyim[I1(j)] = 1.0 / dym[I1(j)]; }
    for (int j = 2; j <= jx - 1; j++) {
        uy1[I1(j)] = 0.5 * dym[I1(j - 1)] / dy[I1(j)];
        uy0[I1(j)] = 0.5 * dym[I1(j)] / dy[I1(j)];
    }
    for (size_t p = 0; p < n; p++) {
        dro[p] = 0.0; dee[p] = 0.0; drx[p] = 0.0; dry[p] = 0.0; drz[p] = 0.0;
        dbx[p] = 0.0; dby[p] = 0.0; dbz[p] = 0.0; daz[p] = 0.0;
    }
    for (size_t p = 0; p < n; p++) { /* :107-115 */
        double vv = sq(vx[p]) + sq(vy[p]) + sq(vz[p]);
        double bb = sq(bx[p]) + sq(by[p]) + sq(bz[p]);
        ee[p] = pr[p] / (gm - 1.0) + 0.5 * ro[p] * vv + pi8i * bb;
        rx[p] = ro[p] * vx[p];
        ry[p] = ro[p] * vy[p];
        rz[p] = ro[p] * vz[p];
    }
    for (size_t p = 0; p < n; p++) { /* :117-123 */
        ex[p] = -vy[p] * bz[p] + vz[p] * by[p];
        ey[p] = -vz[p] * bx[p] + vx[p] * bz[p];
        ez[p] = -vx[p] * by[p] + vy[p] * bx[p];
    }

    /* half step, :129-210 */
    for (size_t p = 0; p < n; p++) { fx[p] = ro[p] * vx[p]; fy[p] = ro[p] * vy[p]; }
    mlwhalf2(ro, roh, dro, dt, fx, dxi, dxim, ix, fy, dyi, dyim, jx);
    for (size_t p = 0; p < n; p++) {
        double vv = sq(vx[p]) + sq(vy[p]) + sq(vz[p]);
        double ep = pr[p] * gm / (gm - 1.0) + 0.5 * ro[p] * vv;
        fx[p] = ep * vx[p] + (bz[p] * ey[p] - by[p] * ez[p]) * pi4i;
        fy[p] = ep * vy[p] + (bx[p] * ez[p] - bz[p] * ex[p]) * pi4i;
    }
    mlwhalf2(ee, eeh, dee, dt, fx, dxi, dxim, ix, fy, dyi, dyim, jx);
    for (size_t p = 0; p < n; p++) {
        fx[p] = ro[p] * sq(vx[p]) + pr[p] + pi8i * (sq(by[p]) + sq(bz[p]) - sq(bx[p]));
        fy[p] = ro[p] * vx[p] * vy[p] - pi4i * bx[p] * by[p];
    }
    mlwhalf2(rx, rxh, drx, dt, fx, dxi, dxim, ix, fy, dyi, dyim, jx);
    for (size_t p = 0; p < n; p++) {
        fx[p] = ro[p] * vy[p] * vx[p] - pi4i * by[p] * bx[p];
        fy[p] = ro[p] * sq(vy[p]) + pr[p] + pi8i * (sq(bx[p]) + sq(bz[p]) - sq(by[p]));
    }
    mlwhalf2(ry, ryh, dry, dt, fx, dxi, dxim, ix, fy, dyi, dyim, jx);
    for (size_t p = 0; p < n; p++) {
        fx[p] = ro[p] * vz[p] * vx[p] - pi4i * bz[p] * bx[p];
        fy[p] = ro[p] * vz[p] * vy[p] - pi4i * bz[p] * by[p];
    }
    mlwhalf2(rz, rzh, drz, dt, fx, dxi, dxim, ix, fy, dyi, dyim, jx);
    for (size_t p = 0; p < n; p++) { fx[p] = 0.0; fy[p] = ez[p]; }
    mlwhalf2(bx, bxh, dbx, dt, fx, dxi, dxim, ix, fy, dyi, dyim, jx);
    for (size_t p = 0; p < n; p++) { fx[p] = -ez[p]; fy[p] = 0.0; }
    mlwhalf2(by, byh, dby, dt, fx, dxi, dxim, ix, fy, dyi, dyim, jx);
    for (size_t p = 0; p < n; p++) { fx[p] = ey[p]; fy[p] = -ex[p]; }
    mlwhalf2(bz, bzh, dbz, dt, fx, dxi, dxim, ix, fy, dyi, dyim, jx);
    for (size_t p = 0; p < n; p++) ss[p] = -ez[p];
    mlwsrch2(azh, daz, dt, ss, ix, jx);

    /* half-step primitives, :214-229 (the E loop runs over every cell in the
     * reference; only corners i<ix, j<jx are ever used) */
    for (int j = 1; j <= jx - 1; j++)
        for (int i = 1; i <= ix - 1; i++) {
            size_t p = I2(i, j);
            vxh[p] = rxh[p] / roh[p];
            vyh[p] = ryh[p] / roh[p];
            vzh[p] = rzh[p] / roh[p];
            double vv = sq(vxh[p]) + sq(vyh[p]) + sq(vzh[p]);
            double bb = sq(bxh[p]) + sq(byh[p]) + sq(bzh[p]);
            prh[p] = (gm - 1.0) * (eeh[p] - 0.5 * roh[p] * vv - pi8i * bb);
            exh[p] = -vyh[p] * bzh[p] + vzh[p] * byh[p];
            eyh[p] = -vzh[p] * bxh[p] + vxh[p] * bzh[p];
            ezh[p] = -vxh[p] * byh[p] + vyh[p] * bxh[p];
        }

    /* full step, :233-318 */
#define HLOOP for (int j = 1; j <= jx - 1; j++) for (int i = 1; i <= ix - 1; i++)
    HLOOP { size_t p = I2(i, j); fx[p] = roh[p] * vxh[p]; fy[p] = roh[p] * vyh[p]; }
    mlwfull2(dro, dt, fx, dxi, ux0, ux1, ix, fy, dyi, uy0, uy1, jx);
    HLOOP {
        size_t p = I2(i, j);
        double vv = sq(vxh[p]) + sq(vyh[p]) + sq(vzh[p]);
        double ep = prh[p] * gm / (gm - 1.0) + 0.5 * roh[p] * vv;
        fx[p] = ep * vxh[p] + (bzh[p] * eyh[p] - byh[p] * ezh[p]) * pi4i;
        fy[p] = ep * vyh[p] + (bxh[p] * ezh[p] - bzh[p] * exh[p]) * pi4i;
    }
    mlwfull2(dee, dt, fx, dxi, ux0, ux1, ix, fy, dyi, uy0, uy1, jx);
    HLOOP {
        size_t p = I2(i, j);
        fx[p] = roh[p] * sq(vxh[p]) + prh[p] + pi8i * (sq(byh[p]) + sq(bzh[p]) - sq(bxh[p]));
        fy[p] = roh[p] * vxh[p] * vyh[p] - pi4i * bxh[p] * byh[p];
    }
    mlwfull2(drx, dt, fx, dxi, ux0, ux1, ix, fy, dyi, uy0, uy1, jx);
    HLOOP {
        size_t p = I2(i, j);
        fx[p] = roh[p] * vyh[p] * vxh[p] - pi4i * byh[p] * bxh[p];
        fy[p] = roh[p] * sq(vyh[p]) + prh[p] + pi8i * (sq(bxh[p]) + sq(bzh[p]) - sq(byh[p]));
    }
    mlwfull2(dry, dt, fx, dxi, ux0, ux1, ix, fy, dyi, uy0, uy1, jx);
    HLOOP {   /* :277 loops over every cell; mlwfull reads only i<ix, j<jx */
        size_t p = I2(i, j);
        fx[p] = roh[p] * vzh[p] * vxh[p] - pi4i * bzh[p] * bxh[p];
        fy[p] = roh[p] * vzh[p] * vyh[p] - pi4i * bzh[p] * byh[p];
    }
    mlwfull2(drz, dt, fx, dxi, ux0, ux1, ix, fy, dyi, uy0, uy1, jx);
    HLOOP { size_t p = I2(i, j); fx[p] = 0.0; fy[p] = ezh[p]; }
    mlwfull2(dbx, dt, fx, dxi, ux0, ux1, ix, fy, dyi, uy0, uy1, jx);
    HLOOP { size_t p = I2(i, j); fx[p] = -ezh[p]; fy[p] = 0.0; }
    mlwfull2(dby, dt, fx, dxi, ux0, ux1, ix, fy, dyi, uy0, uy1, jx);
    HLOOP { size_t p = I2(i, j); fx[p] = eyh[p]; fy[p] = -exh[p]; }
    mlwfull2(dbz, dt, fx, dxi, ux0, ux1, ix, fy, dyi, uy0, uy1, jx);
    HLOOP { size_t p = I2(i, j); ss[p] = -ezh[p]; }
#undef HLOOP
    mlwsrcf2(daz, dt, ss, ux0, ux1, ix, uy0, uy1, jx);

    /* artificial viscosity, :324-345 */
    const double zero = 0.0;
    for (int j = 1; j <= jx - 1; j++)
        for (int i = 1; i <= ix - 1; i++)
            qx[I2(i, j)] = qav * dxm[I1(i)] * fmaxF(zero, fabs(vx[I2(i + 1, j)] - vx[I2(i, j)]) - R4_1EM4);
    for (int j = 1; j <= jx - 1; j++)
        for (int i = 1; i <= ix; i++)
            qy[I2(i, j)] = qav * dym[I1(j)] * fmaxF(zero, fabs(vy[I2(i, j + 1)] - vy[I2(i, j)]) - R4_1EM4);
    mlwartv2(ro, dro, dt, qx, dxi, dxim, ix, qy, dyi, dyim, jx);
    mlwartv2(ee, dee, dt, qx, dxi, dxim, ix, qy, dyi, dyim, jx);
    mlwartv2(rx, drx, dt, qx, dxi, dxim, ix, qy, dyi, dyim, jx);
    mlwartv2(ry, dry, dt, qx, dxi, dxim, ix, qy, dyi, dyim, jx);
    mlwartv2(rz, drz, dt, qx, dxi, dxim, ix, qy, dyi, dyim, jx);
    mlwartv2(bx, dbx, dt, qx, dxi, dxim, ix, qy, dyi, dyim, jx);
    mlwartv2(by, dby, dt, qx, dxi, dxim, ix, qy, dyi, dyim, jx);
    mlwartv2(bz, dbz, dt, qx, dxi, dxim, ix, qy, dyi, dyim, jx);

    /* update and primitive recovery, :349-374 */
    for (int j = 2; j <= jx - 1; j++)
        for (int i = 2; i <= ix - 1; i++) {
            size_t p = I2(i, j);
            ro[p] = ro[p] + dro[p];
            ee[p] = ee[p] + dee[p];
            rx[p] = rx[p] + drx[p];
            ry[p] = ry[p] + dry[p];
            rz[p] = rz[p] + drz[p];
            bx[p] = bx[p] + dbx[p];
            by[p] = by[p] + dby[p];
            bz[p] = bz[p] + dbz[p];
            az[p] = az[p] + daz[p];
        }
    for (int j = 2; j <= jx - 1; j++)
        for (int i = 2; i <= ix - 1; i++) {
            size_t p = I2(i, j);
            vx[p] = rx[p] / ro[p];
            vy[p] = ry[p] / ro[p];
            vz[p] = rz[p] / ro[p];
            double vv = sq(vx[p]) + sq(vy[p]) + sq(vz[p]);
            double bb = sq(bx[p]) + sq(by[p]) + sq(bz[p]);
            pr[p] = (gm - 1.0) * (ee[p] - 0.5 * ro[p] * vv - pi8i * bb);
        }
    free(dxi); free(dxim); free(ux0); free(ux1); free(dyi); free(dyim); free(uy0); free(uy1);
    free(ee); free(rx); free(ry); free(rz); free(ex); free(ey); free(ez);
    free(roh); free(eeh); free(rxh); free(ryh); free(rzh); free(bxh); free(byh); free(bzh); free(azh);
    free(prh); free(vxh); free(vyh); free(vzh); free(exh); free(eyh); free(ezh);
    free(dro); free(dee); free(drx); free(dry); free(drz); free(dbx); free(dby); free(dbz); free(daz);
    free(fx); free(fy); free(qx); free(qy); free(ss);
}

/* common/cfl_m3.f:42-78: over every cell, min(dx,dy)/sqrt(v^2+cs^2+ca^2) */
void or2d_cfl_m3(double *dt, double safety, double dtmin, int *merr, double gm,
                 const double *ro, const double *pr, const double *vx,
                 const double *vy, const double *vz, const double *bx,
                 const double *by, const double *bz, const double *dx, int ix,
                 const double *dy, int jx, int *imin, int *jmin) {
    const double pi = acos(-1.0);
    const double pi4i = 1.0 / pi / 4.0;
    size_t n = (size_t)ix * jx;
    double *dtq = dalloc(n);
    for (int j = 1; j <= jx; j++)
        for (int i = 1; i <= ix; i++) {
            size_t p = I2(i, j);
            double b2 = sq(bx[p]) + sq(by[p]) + sq(bz[p]);
            double ca2 = b2 * pi4i / ro[p];
            double cs2 = gm * pr[p] / ro[p];
            double v2 = vx[p] * vx[p] + vy[p] * vy[p] + sq(vz[p]);
            double dtcfl = fminF(dx[I1(i)], dy[I1(j)]) / sqrt(v2 + cs2 + ca2);
            dtq[p] = safety * dtcfl;
        }
    double d = R4_1E20;
    int im = 0, jm = 0;
    for (int j = 1; j <= jx; j++)
        for (int i = 1; i <= ix; i++)
            if (dtq[I2(i, j)] < d) { im = i; jm = j; d = dtq[I2(i, j)]; }
    *merr = (d < dtmin) ? 9001 : 0;
    *dt = d;
    if (imin) *imin = im;
    if (jmin) *jmin = jm;
    free(dtq);
}

/* hdmlw/mlw_m_c.f:2-350: mlw_m in axisymmetric cylindrical (r, z) = (x, z)
 * coordinates; fields ro, pr, vx, vz, bx, bz and the potential Ay.  Every
 * variable but bx carries the geometric source -F_x/r through mlwsrch (cell,
 * r = x) and mlwsrcf (corner, r = xm; for bz the reference divides by x). */
void or2d_mlw_m_c(double *ro, double *pr, double *vx, double *vz, double *bx,
                  double *bz, double *ay, double dt, double qav, double gm,
                  const double *x, const double *xm, const double *dx,
                  const double *dxm, int ix, const double *dz,
                  const double *dzm, int jx) {
    const double pi = acos(-1.0);
    const double pi8i = 1.0 / pi / 8.0;
    const double pi4i = 1.0 / pi / 4.0;
    size_t n = (size_t)ix * jx;
    double *dxi = dalloc(ix), *dxim = dalloc(ix), *ux0 = dalloc(ix), *ux1 = dalloc(ix);
    double *dzi = dalloc(jx), *dzim = dalloc(jx), *uz0 = dalloc(jx), *uz1 = dalloc(jx);
    double *ee = dalloc(n), *rx = dalloc(n), *rz = dalloc(n), *ey = dalloc(n);
    double *roh = dalloc(n), *eeh = dalloc(n), *rxh = dalloc(n), *rzh = dalloc(n), *bxh = dalloc(n), *bzh = dalloc(n);
    double *ayh = dalloc(n), *eyh = dalloc(n), *prh = dalloc(n), *vxh = dalloc(n), *vzh = dalloc(n);
    double *dro = dalloc(n), *dee = dalloc(n), *drx = dalloc(n), *drz = dalloc(n), *dbx = dalloc(n), *dbz = dalloc(n);
    double *day = dalloc(n), *fx = dalloc(n), *fz = dalloc(n), *qx = dalloc(n), *qz = dalloc(n), *ss = dalloc(n);

    for (int i = 1; i <= ix; i++) { dxi[I1(i)] = 1.0 / dx[I1(i)]; dxim[I1(i)] = 1.0 / dxm[I1(i)]; } /* :70-73 */
    for (int i = 2; i <= ix - 1; i++) {
        ux1[I1(i)] = 0.5 * dxm[I1(i - 1)] / dx[I1(i)];
        ux0[I1(i)] = 0.5 * dxm[I1(i)] / dx[I1(i)];
    }
    for (int j = 1; j <= jx; j++) { dzi[I1(j)] = 1.0 / dz[I1(j)]; dzim[I1(j)] = 1.0 / dzm[I1(j)]; }
    for (int j = 2; j <= jx - 1; j++) {
        uz1[I1(j)] = 0.5 * dzm[I1(j - 1)] / dz[I1(j)];
        uz0[I1(j)] = 0.5 * dzm[I1(j)] / dz[I1(j)];
    }
    for (size_t p = 0; p < n; p++) {
        dro[p] = 0.0; dee[p] = 0.0; drx[p] = 0.0; drz[p] = 0.0; dbx[p] = 0.0; dbz[p] = 0.0; day[p] = 0.0;
    }
#define ALL for (int j = 1; j <= jx; j++) for (int i = 1; i <= ix; i++)
#define HLF for (int j = 1; j <= jx - 1; j++) for (int i = 1; i <= ix - 1; i++)
    ALL { /* :105-112 */
        size_t p = I2(i, j);
        double vv = sq(vx[p]) + sq(vz[p]);
        double bb = sq(bx[p]) + sq(bz[p]);
        ee[p] = pr[p] / (gm - 1.0) + 0.5 * ro[p] * vv + pi8i * bb;
        rx[p] = ro[p] * vx[p];
        rz[p] = ro[p] * vz[p];
    }
    ALL { size_t p = I2(i, j); ey[p] = -vz[p] * bx[p] + vx[p] * bz[p]; }

    /* half step with the cell sources, :123-198 */
    ALL { size_t p = I2(i, j); fx[p] = ro[p] * vx[p]; fz[p] = ro[p] * vz[p]; ss[p] = -fx[p] / x[I1(i)]; }
    mlwhalf2(ro, roh, dro, dt, fx, dxi, dxim, ix, fz, dzi, dzim, jx);
    mlwsrch2(roh, dro, dt, ss, ix, jx);
    ALL {
        size_t p = I2(i, j);
        double vv = sq(vx[p]) + sq(vz[p]);
        double ep = pr[p] * gm / (gm - 1.0) + 0.5 * ro[p] * vv;
        fx[p] = ep * vx[p] + (bz[p] * ey[p]) * pi4i;
        fz[p] = ep * vz[p] + (-bx[p] * ey[p]) * pi4i;
        ss[p] = -fx[p] / x[I1(i)];
    }
    mlwhalf2(ee, eeh, dee, dt, fx, dxi, dxim, ix, fz, dzi, dzim, jx);
    mlwsrch2(eeh, dee, dt, ss, ix, jx);
    ALL {
        size_t p = I2(i, j);
        fx[p] = ro[p] * sq(vx[p]) + pr[p] + pi8i * (sq(bz[p]) - sq(bx[p]));
        fz[p] = ro[p] * vx[p] * vz[p] - pi4i * bx[p] * bz[p];
        ss[p] = -(ro[p] * sq(vx[p]) + pi4i * (-sq(bx[p]))) / x[I1(i)];
    }
    mlwhalf2(rx, rxh, drx, dt, fx, dxi, dxim, ix, fz, dzi, dzim, jx);
    mlwsrch2(rxh, drx, dt, ss, ix, jx);
    ALL {
        size_t p = I2(i, j);
        fx[p] = ro[p] * vz[p] * vx[p] - pi4i * bz[p] * bx[p];
        fz[p] = ro[p] * sq(vz[p]) + pr[p] + pi8i * (sq(bx[p]) - sq(bz[p]));
        ss[p] = -fx[p] / x[I1(i)];
    }
    mlwhalf2(rz, rzh, drz, dt, fx, dxi, dxim, ix, fz, dzi, dzim, jx);
    mlwsrch2(rzh, drz, dt, ss, ix, jx);
    ALL { size_t p = I2(i, j); fx[p] = 0.0; fz[p] = -ey[p]; }
    mlwhalf2(bx, bxh, dbx, dt, fx, dxi, dxim, ix, fz, dzi, dzim, jx);
    ALL { size_t p = I2(i, j); fx[p] = ey[p]; fz[p] = 0.0; ss[p] = -fx[p] / x[I1(i)]; }
    mlwhalf2(bz, bzh, dbz, dt, fx, dxi, dxim, ix, fz, dzi, dzim, jx);
    mlwsrch2(bzh, dbz, dt, ss, ix, jx);
    ALL { size_t p = I2(i, j); ss[p] = -ey[p]; }
    mlwsrch2(ayh, day, dt, ss, ix, jx);

    /* half-step primitives, :202-215 (E over the corners that are used) */
    HLF {
        size_t p = I2(i, j);
        vxh[p] = rxh[p] / roh[p];
        vzh[p] = rzh[p] / roh[p];
        double vv = sq(vxh[p]) + sq(vzh[p]);
        double bb = sq(bxh[p]) + sq(bzh[p]);
        prh[p] = (gm - 1.0) * (eeh[p] - 0.5 * roh[p] * vv - pi8i * bb);
        eyh[p] = -vzh[p] * bxh[p] + vxh[p] * bzh[p];
    }

    /* full step with the corner sources, :219-296 */
    HLF { size_t p = I2(i, j); fx[p] = roh[p] * vxh[p]; fz[p] = roh[p] * vzh[p]; ss[p] = -fx[p] / xm[I1(i)]; }
    mlwfull2(dro, dt, fx, dxi, ux0, ux1, ix, fz, dzi, uz0, uz1, jx);
    mlwsrcf2(dro, dt, ss, ux0, ux1, ix, uz0, uz1, jx);
    HLF {
        size_t p = I2(i, j);
        double vv = sq(vxh[p]) + sq(vzh[p]);
        double ep = prh[p] * gm / (gm - 1.0) + 0.5 * roh[p] * vv;
        fx[p] = ep * vxh[p] + (bzh[p] * eyh[p]) * pi4i;
        fz[p] = ep * vzh[p] + (-bxh[p] * eyh[p]) * pi4i;
        ss[p] = -fx[p] / xm[I1(i)];
    }
    mlwfull2(dee, dt, fx, dxi, ux0, ux1, ix, fz, dzi, uz0, uz1, jx);
    mlwsrcf2(dee, dt, ss, ux0, ux1, ix, uz0, uz1, jx);
    HLF {
        size_t p = I2(i, j);
        fx[p] = roh[p] * sq(vxh[p]) + prh[p] + pi8i * (sq(bzh[p]) - sq(bxh[p]));
        fz[p] = roh[p] * vxh[p] * vzh[p] - pi4i * bxh[p] * bzh[p];
        ss[p] = -(roh[p] * sq(vxh[p]) + pi4i * (-sq(bxh[p]))) / xm[I1(i)];
    }
    mlwfull2(drx, dt, fx, dxi, ux0, ux1, ix, fz, dzi, uz0, uz1, jx);
    mlwsrcf2(drx, dt, ss, ux0, ux1, ix, uz0, uz1, jx);
    HLF {
        size_t p = I2(i, j);
        fx[p] = roh[p] * vzh[p] * vxh[p] - pi4i * bzh[p] * bxh[p];
        fz[p] = roh[p] * sq(vzh[p]) + prh[p] + pi8i * (sq(bxh[p]) - sq(bzh[p]));
        ss[p] = -fx[p] / xm[I1(i)];
    }
    mlwfull2(drz, dt, fx, dxi, ux0, ux1, ix, fz, dzi, uz0, uz1, jx);
    mlwsrcf2(drz, dt, ss, ux0, ux1, ix, uz0, uz1, jx);
    HLF { size_t p = I2(i, j); fx[p] = 0.0; fz[p] = -eyh[p]; }
    mlwfull2(dbx, dt, fx, dxi, ux0, ux1, ix, fz, dzi, uz0, uz1, jx);
    HLF { size_t p = I2(i, j); fx[p] = eyh[p]; fz[p] = 0.0; ss[p] = -fx[p] / x[I1(i)]; }   /* :284: x, not xm */
    mlwfull2(dbz, dt, fx, dxi, ux0, ux1, ix, fz, dzi, uz0, uz1, jx);
    mlwsrcf2(dbz, dt, ss, ux0, ux1, ix, uz0, uz1, jx);
    HLF { size_t p = I2(i, j); ss[p] = -eyh[p]; }
    mlwsrcf2(day, dt, ss, ux0, ux1, ix, uz0, uz1, jx);
#undef ALL
#undef HLF

    /* artificial viscosity, :302-321 */
    const double zero = 0.0;
    for (int j = 1; j <= jx - 1; j++)
        for (int i = 1; i <= ix - 1; i++)
            qx[I2(i, j)] = qav * dxm[I1(i)] * fmaxF(zero, fabs(vx[I2(i + 1, j)] - vx[I2(i, j)]) - R4_1EM4);
    for (int j = 1; j <= jx - 1; j++)
        for (int i = 1; i <= ix; i++)
            qz[I2(i, j)] = qav * dzm[I1(j)] * fmaxF(zero, fabs(vz[I2(i, j + 1)] - vz[I2(i, j)]) - R4_1EM4);
    mlwartv2(ro, dro, dt, qx, dxi, dxim, ix, qz, dzi, dzim, jx);
    mlwartv2(ee, dee, dt, qx, dxi, dxim, ix, qz, dzi, dzim, jx);
    mlwartv2(rx, drx, dt, qx, dxi, dxim, ix, qz, dzi, dzim, jx);
    mlwartv2(rz, drz, dt, qx, dxi, dxim, ix, qz, dzi, dzim, jx);
    mlwartv2(bx, dbx, dt, qx, dxi, dxim, ix, qz, dzi, dzim, jx);
    mlwartv2(bz, dbz, dt, qx, dxi, dxim, ix, qz, dzi, dzim, jx);

    /* update and primitive recovery, :325-346 */
    for (int j = 2; j <= jx - 1; j++)
        for (int i = 2; i <= ix - 1; i++) {
            size_t p = I2(i, j);
            ro[p] = ro[p] + dro[p];
            ee[p] = ee[p] + dee[p];
            rx[p] = rx[p] + drx[p];
            rz[p] = rz[p] + drz[p];
            bx[p] = bx[p] + dbx[p];
            bz[p] = bz[p] + dbz[p];
            ay[p] = ay[p] + day[p];
        }
    for (int j = 2; j <= jx - 1; j++)
        for (int i = 2; i <= ix - 1; i++) {
            size_t p = I2(i, j);
            vx[p] = rx[p] / ro[p];
            vz[p] = rz[p] / ro[p];
            double vv = sq(vx[p]) + sq(vz[p]);
            double bb = sq(bx[p]) + sq(bz[p]);
            pr[p] = (gm - 1.0) * (ee[p] - 0.5 * ro[p] * vv - pi8i * bb);
        }
    double *tofree[] = {dxi, dxim, ux0, ux1, dzi, dzim, uz0, uz1, ee, rx, rz, ey, roh, eeh, rxh, rzh, bxh, bzh,
                        ayh, eyh, prh, vxh, vzh, dro, dee, drx, drz, dbx, dbz, day, fx, fz, qx, qz, ss};
    for (size_t t = 0; t < sizeof(tofree) / sizeof(tofree[0]); t++) free(tofree[t]);
}

/* hdmlw/mlw_m_s.f:1-398 (2D spherical (r, theta) MHD with Az): the conserved
 * variables and fluxes carry the metric factors, rx and ry the geometric
 * sources; the ry x flux at the corners carries sin(ym) alone
 * (mlw_m_s.f:305), kept as in the reference.  sin, cos of y and ym from the
 * C library, as the reference's compiled intrinsics. */
void or2d_mlw_m_s(double *ro, double *pr, double *vx, double *vy, double *bx, double *by, double *az,
                  double dt, double qav, double gm, const double *x, const double *xm, const double *dx,
                  const double *dxm, int ix, const double *y, const double *ym, const double *dy,
                  const double *dym, int jx) {
    const double pi = acos(-1.0);
    const double pi8i = 1.0 / pi / 8.0;
    const double pi4i = 1.0 / pi / 4.0;
    size_t n = (size_t)ix * jx;
    double *dxi = dalloc(ix), *dxim = dalloc(ix), *ux0 = dalloc(ix), *ux1 = dalloc(ix);
    double *dyi = dalloc(jx), *dyim = dalloc(jx), *uy0 = dalloc(jx), *uy1 = dalloc(jx);
    double *siny = dalloc(jx), *sinym = dalloc(jx);
    double *ee = dalloc(n), *rx = dalloc(n), *ry = dalloc(n), *ez = dalloc(n);
    double *sc[6], *sch[6], *dsc[6];
    for (int q = 0; q < 6; q++) { sc[q] = dalloc(n); sch[q] = dalloc(n); dsc[q] = dalloc(n); }
    double *daz = dalloc(n), *azh = dalloc(n);
    double *roh = dalloc(n), *eeh = dalloc(n), *prh = dalloc(n), *vxh = dalloc(n), *vyh = dalloc(n);
    double *bxh = dalloc(n), *byh = dalloc(n), *ezh = dalloc(n);
    double *fx = dalloc(n), *fy = dalloc(n), *ss = dalloc(n), *qx = dalloc(n), *qy = dalloc(n);
    for (int i = 1; i <= ix; i++) { dxi[I1(i)] = 1.0 / dx[I1(i)]; dxim[I1(i)] = 1.0 / dxm[I1(i)]; }
    for (int i = 2; i <= ix - 1; i++) { ux1[I1(i)] = 0.5 * dxm[I1(i - 1)] / dx[I1(i)]; ux0[I1(i)] = 0.5 * dxm[I1(i)] / dx[I1(i)]; }
    for (int j = 1; j <= jx; j++) { dyi[I1(j)] = 1.0 / dy[I1(j)]; dyim[I1(j)] = 1.0 / dym[I1(j)]; }
    for (int j = 2; j <= jx - 1; j++) { uy1[I1(j)] = 0.5 * dym[I1(j - 1)] / dy[I1(j)]; uy0[I1(j)] = 0.5 * dym[I1(j)] / dy[I1(j)]; }
    for (int j = 1; j <= jx; j++) { siny[I1(j)] = sin(y[I1(j)]); sinym[I1(j)] = sin(ym[I1(j)]); }   /* :99-102 */
    for (size_t p = 0; p < n; p++) {   /* :107-133 */
        for (int q = 0; q < 6; q++) dsc[q][p] = 0.0;
        daz[p] = 0.0; azh[p] = 0.0;
        double vv = sq(vx[p]) + sq(vy[p]);
        double bb = sq(bx[p]) + sq(by[p]);
        ee[p] = pr[p] / (gm - 1.0) + 0.5 * ro[p] * vv + pi8i * bb;
        rx[p] = ro[p] * vx[p];
        ry[p] = ro[p] * vy[p];
        ez[p] = -vx[p] * by[p] + vy[p] * bx[p];
    }
#define CELLS for (int j = 1; j <= jx; j++) for (int i = 1; i <= ix; i++)
#define HALF(q) mlwhalf2(sc[q], sch[q], dsc[q], dt, fx, dxi, dxim, ix, fy, dyi, dyim, jx)
    CELLS { /* :136-145 */
        size_t p = I2(i, j); double r = x[I1(i)], st = siny[I1(j)];
        sc[0][p] = ro[p] * sq(r) * st; sc[1][p] = ee[p] * sq(r) * st; sc[2][p] = rx[p] * sq(r) * st;
        sc[3][p] = ry[p] * r * st;      sc[4][p] = bx[p] * sq(r) * st; sc[5][p] = by[p] * r * st;
    }
    CELLS { size_t p = I2(i, j); double r = x[I1(i)], st = siny[I1(j)];   /* :148-157 */
            fx[p] = ro[p] * vx[p]; fy[p] = ro[p] * vy[p];
            fx[p] = fx[p] * sq(r) * st; fy[p] = fy[p] * r * st; }
    HALF(0);
    CELLS { /* :160-170 */
        size_t p = I2(i, j); double r = x[I1(i)], st = siny[I1(j)];
        double vv = sq(vx[p]) + sq(vy[p]);
        double ep = pr[p] * gm / (gm - 1.0) + 0.5 * ro[p] * vv;
        fx[p] = ep * vx[p] - by[p] * ez[p] * pi4i;
        fy[p] = ep * vy[p] + bx[p] * ez[p] * pi4i;
        fx[p] = fx[p] * sq(r) * st; fy[p] = fy[p] * r * st;
    }
    HALF(1);
    CELLS { /* :173-184 */
        size_t p = I2(i, j); double r = x[I1(i)], st = siny[I1(j)];
        fx[p] = ro[p] * sq(vx[p]) + pr[p] + pi8i * (sq(by[p]) - sq(bx[p]));
        fy[p] = ro[p] * vx[p] * vy[p] - pi4i * bx[p] * by[p];
        fx[p] = fx[p] * sq(r) * st; fy[p] = fy[p] * r * st;
        ss[p] = (ro[p] * sq(vy[p]) + 2.0 * pr[p] + pi4i * sq(bx[p])) * r * st;
    }
    HALF(2); mlwsrch2(sch[2], dsc[2], dt, ss, ix, jx);
    CELLS { /* :188-201 */
        size_t p = I2(i, j); double r = x[I1(i)], st = siny[I1(j)];
        fx[p] = ro[p] * vy[p] * vx[p] - pi4i * by[p] * bx[p];
        fy[p] = ro[p] * sq(vy[p]) + pr[p] + pi8i * (sq(bx[p]) - sq(by[p]));
        fx[p] = fx[p] * r * st; fy[p] = fy[p] * st;
        ss[p] = (ro[p] * vx[p] * vy[p] - pi4i * bx[p] * by[p]) * (-2.0) * st
              + (pr[p] + pi8i * (sq(bx[p]) + sq(by[p]))) * cos(y[I1(j)]);
    }
    HALF(3); mlwsrch2(sch[3], dsc[3], dt, ss, ix, jx);
    CELLS { size_t p = I2(i, j); fx[p] = 0.0; fy[p] = ez[p]; fy[p] = fy[p] * x[I1(i)] * siny[I1(j)]; }   /* :206-212 */
    HALF(4);
    CELLS { size_t p = I2(i, j); fx[p] = -ez[p]; fy[p] = 0.0; fx[p] = fx[p] * x[I1(i)] * siny[I1(j)]; }  /* :216-222 */
    HALF(5);
    CELLS { size_t p = I2(i, j); ss[p] = -ez[p]; }                                                      /* :226-231 */
    mlwsrch2(azh, daz, dt, ss, ix, jx);
#undef HALF
#undef CELLS
    for (int j = 1; j <= jx - 1; j++) /* :234-247 */
        for (int i = 1; i <= ix - 1; i++) {
            size_t p = I2(i, j); double r = xm[I1(i)], st = sinym[I1(j)];
            roh[p] = sch[0][p] / sq(r) / st;
            eeh[p] = sch[1][p] / sq(r) / st;
            vxh[p] = sch[2][p] / roh[p] / sq(r) / st;
            vyh[p] = sch[3][p] / roh[p] / r / st;
            bxh[p] = sch[4][p] / sq(r) / st;
            byh[p] = sch[5][p] / r / st;
            double vv = sq(vxh[p]) + sq(vyh[p]);
            double bb = sq(bxh[p]) + sq(byh[p]);
            prh[p] = (gm - 1.0) * (eeh[p] - 0.5 * roh[p] * vv - pi8i * bb);
            ezh[p] = -vxh[p] * byh[p] + vyh[p] * bxh[p];
        }
#define FULL(q) mlwfull2(dsc[q], dt, fx, dxi, ux0, ux1, ix, fy, dyi, uy0, uy1, jx)
#define SRCF(du) mlwsrcf2(du, dt, ss, ux0, ux1, ix, uy0, uy1, jx)
#define HL for (int j = 1; j <= jx - 1; j++) for (int i = 1; i <= ix - 1; i++)
    HL { size_t p = I2(i, j); double r = xm[I1(i)], st = sinym[I1(j)];   /* :258-265 */
         fx[p] = roh[p] * vxh[p]; fy[p] = roh[p] * vyh[p];
         fx[p] = fx[p] * sq(r) * st; fy[p] = fy[p] * r * st; }
    FULL(0);
    HL { /* :268-278 */
        size_t p = I2(i, j); double r = xm[I1(i)], st = sinym[I1(j)];
        double vv = sq(vxh[p]) + sq(vyh[p]);
        double eph = prh[p] * gm / (gm - 1.0) + 0.5 * roh[p] * vv;
        fx[p] = eph * vxh[p] - byh[p] * ezh[p] * pi4i;
        fy[p] = eph * vyh[p] + bxh[p] * ezh[p] * pi4i;
        fx[p] = fx[p] * sq(r) * st; fy[p] = fy[p] * r * st;
    }
    FULL(1);
    HL { /* :282-293 */
        size_t p = I2(i, j); double r = xm[I1(i)], st = sinym[I1(j)];
        fx[p] = roh[p] * sq(vxh[p]) + prh[p] + pi8i * (sq(byh[p]) - sq(bxh[p]));
        fy[p] = roh[p] * vxh[p] * vyh[p] - pi4i * bxh[p] * byh[p];
        fx[p] = fx[p] * sq(r) * st; fy[p] = fy[p] * r * st;
        ss[p] = (roh[p] * sq(vyh[p]) + 2.0 * prh[p] + pi4i * sq(bxh[p])) * r * st;
    }
    FULL(2); SRCF(dsc[2]);
    HL { /* :298-310 */
        size_t p = I2(i, j); double st = sinym[I1(j)];
        fx[p] = roh[p] * vyh[p] * vxh[p] - pi4i * byh[p] * bxh[p];
        fy[p] = roh[p] * sq(vyh[p]) + prh[p] + pi8i * (sq(bxh[p]) - sq(byh[p]));
        fx[p] = fx[p] * st; fy[p] = fy[p] * st;
        ss[p] = (roh[p] * vxh[p] * vyh[p] - pi4i * bxh[p] * byh[p]) * (-2.0) * st
              + (prh[p] + pi8i * (sq(bxh[p]) + sq(byh[p]))) * cos(ym[I1(j)]);
    }
    FULL(3); SRCF(dsc[3]);
    HL { size_t p = I2(i, j); fx[p] = 0.0; fy[p] = ezh[p]; fy[p] = fy[p] * xm[I1(i)] * sinym[I1(j)]; }   /* :315-321 */
    FULL(4);
    HL { size_t p = I2(i, j); fx[p] = -ezh[p]; fy[p] = 0.0; fx[p] = fx[p] * xm[I1(i)] * sinym[I1(j)]; }  /* :325-331 */
    FULL(5);
    HL { size_t p = I2(i, j); ss[p] = -ezh[p]; }                                                        /* :335-340 */
    SRCF(daz);
#undef HL
#undef SRCF
#undef FULL
    const double zero = 0.0;   /* :346-357 */
    for (int j = 1; j <= jx; j++)
        for (int i = 1; i <= ix - 1; i++)
            qx[I2(i, j)] = qav * dxm[I1(i)] * fmaxF(zero, fabs(vx[I2(i + 1, j)] - vx[I2(i, j)]) - R4_1EM4);
    for (int j = 1; j <= jx - 1; j++)
        for (int i = 1; i <= ix; i++)
            qy[I2(i, j)] = qav * dym[I1(j)] * fmaxF(zero, fabs(vy[I2(i, j + 1)] - vy[I2(i, j)]) - R4_1EM4);
    for (int q = 0; q < 6; q++) mlwartv2(sc[q], dsc[q], dt, qx, dxi, dxim, ix, qy, dyi, dyim, jx);   /* :360-365 */
    for (int j = 2; j <= jx - 1; j++) /* :369-396 */
        for (int i = 2; i <= ix - 1; i++) {
            size_t p = I2(i, j); double r = x[I1(i)], st = siny[I1(j)];
            double u[6];
            for (int q = 0; q < 6; q++) u[q] = sc[q][p] + dsc[q][p];
            az[p] = az[p] + daz[p];
            ro[p] = u[0] / sq(r) / st;
            double e = u[1] / sq(r) / st;
            vx[p] = u[2] / ro[p] / sq(r) / st;
            vy[p] = u[3] / ro[p] / r / st;
            bx[p] = u[4] / sq(r) / st;
            by[p] = u[5] / r / st;
            double vv = sq(vx[p]) + sq(vy[p]);
            double bb = sq(bx[p]) + sq(by[p]);
            pr[p] = (gm - 1.0) * (e - 0.5 * ro[p] * vv - pi8i * bb);
        }
    double *tofree[] = {dxi, dxim, ux0, ux1, dyi, dyim, uy0, uy1, siny, sinym, ee, rx, ry, ez, daz, azh, roh, eeh, prh,
                        vxh, vyh, bxh, byh, ezh, fx, fy, ss, qx, qy};
    for (size_t t = 0; t < sizeof(tofree) / sizeof(tofree[0]); t++) free(tofree[t]);
    for (int q = 0; q < 6; q++) { free(sc[q]); free(sch[q]); free(dsc[q]); }
}

/* common/cfl_m_s.f:1-71: cfl_m with the widths dx, x dy (pi in double) */
void or2d_cfl_m_s(double *dt, double safety, double dtmin, int *merr, double gm,
                  const double *ro, const double *pr, const double *vx, const double *vy, const double *bx,
                  const double *by, const double *x, const double *dx, int ix, const double *dy, int jx,
                  int *imin, int *jmin) {
    const double pi = acos(-1.0);
    const double pi4i = 1.0 / pi / 4.0;
    double d = R4_1E20;
    int im = 0, jm = 0;
    for (int j = 1; j <= jx; j++)   /* :48-67, the two loops fused in the same order */
        for (int i = 1; i <= ix; i++) {
            size_t p = I2(i, j);
            double b2 = bx[p] * bx[p] + by[p] * by[p];
            double ca2 = b2 * pi4i / ro[p];
            double cs2 = gm * pr[p] / ro[p];
            double v2 = vx[p] * vx[p] + vy[p] * vy[p];
            double q = safety * (fminF(dx[I1(i)], x[I1(i)] * dy[I1(j)]) / sqrt(v2 + cs2 + ca2));
            if (q < d) { im = i; jm = j; d = q; }
        }
    *merr = (d < dtmin) ? 9001 : 0;
    *dt = d;
    if (imin) *imin = im;
    if (jmin) *jmin = jm;
}

/* ---- the helper routines themselves, with the reference argument lists */
void or2d_mlwhalf(const double *u, double *un, double *du, double dt, const double *fx, const double *dxi,
                  const double *dxim, int ix, const double *fy, const double *dyi, const double *dyim, int jx) {
    mlwhalf2(u, un, du, dt, fx, dxi, dxim, ix, fy, dyi, dyim, jx);
}
void or2d_mlwfull(double *du, double dt, const double *fx, const double *dxi, const double *ux0, const double *ux1,
                  int ix, const double *fy, const double *dyi, const double *uy0, const double *uy1, int jx) {
    mlwfull2(du, dt, fx, dxi, ux0, ux1, ix, fy, dyi, uy0, uy1, jx);
}
void or2d_mlwartv(const double *u, double *du, double dt, const double *qx, const double *dxi, const double *dxim,
                  int ix, const double *qy, const double *dyi, const double *dyim, int jx) {
    mlwartv2(u, du, dt, qx, dxi, dxim, ix, qy, dyi, dyim, jx);
}
void or2d_mlwsrch(double *un, double *du, double dt, const double *s, int ix, int jx) {
    mlwsrch2(un, du, dt, s, ix, jx);
}
void or2d_mlwsrcf(double *du, double dt, const double *s, const double *ux0, const double *ux1, int ix,
                  const double *uy0, const double *uy1, int jx) {
    mlwsrcf2(du, dt, s, ux0, ux1, ix, uy0, uy1, jx);
}
