/*
 * cans_oracle1d.c -- C restatement of the cans1d hot path (test
 * infrastructure only; see cans_oracle.h).  Citations are to
 * cans_beta_vary/cans1d/.
 */
#include "cans_oracle.h"
#include "cans_oracle_util.h"

/* hdmlw/mlwhalf.f:33-46 */
static void mlwhalf1(const double *u, double *un, double *du, double dt,
                     const double *f, const double *dxi, const double *dxim,
                     int ix) {
    for (int i = 2; i <= ix - 1; i++)
        du[I1(i)] = du[I1(i)] - 0.5 * dt * (0.5 * dxi[I1(i)] * (f[I1(i + 1)] - f[I1(i - 1)]));
    for (int i = 1; i <= ix - 1; i++) {
        double uh = 0.5 * (u[I1(i + 1)] + u[I1(i)]);
        double dfdx = dxim[I1(i)] * (f[I1(i + 1)] - f[I1(i)]);
        un[I1(i)] = uh - dt * dfdx;
    }
}

/* hdmlw/mlwfull.f:30-33 */
static void mlwfull1(double *du, double dt, const double *f, const double *dxi,
                     int ix) {
    for (int i = 2; i <= ix - 1; i++) {
        double dfdx = dxi[I1(i)] * (f[I1(i)] - f[I1(i - 1)]);
        du[I1(i)] = du[I1(i)] - 0.5 * dt * dfdx;
    }
}

/* hdmlw/mlwartv.f:31-34 */
static void mlwartv1(const double *u, double *du, double dt, const double *qx,
                     const double *dxi, const double *dxim, int ix) {
    for (int i = 2; i <= ix - 1; i++)
        du[I1(i)] = du[I1(i)] + dt * (dxi[I1(i)] * (qx[I1(i)] * dxim[I1(i)] * (u[I1(i + 1)] - u[I1(i)])
                                                    - qx[I1(i - 1)] * dxim[I1(i - 1)] * (u[I1(i)] - u[I1(i - 1)])));
}

/* hdmlw/mlw_m.f:2-215 (modified Lax-Wendroff, 1D MHD, Gaussian units) */
void or1d_mlw_m(double *ro, double *pr, double *vx, double *vy, double *by,
                const double *bx, const double *bxm, double dt, double qav,
                double gm, const double *dx, const double *dxm, int ix) {
    const double pi = acos(-1.0);
    const double pi8i = 1.0 / pi / 8.0; /* mlw_m.f:55 */
    const double pi4i = 1.0 / pi / 4.0; /* mlw_m.f:56 */
    double *dxi = dalloc(ix), *dxim = dalloc(ix);
    double *ee = dalloc(ix), *rx = dalloc(ix), *ry = dalloc(ix), *ez = dalloc(ix);
    double *roh = dalloc(ix), *eeh = dalloc(ix), *rxh = dalloc(ix), *ryh = dalloc(ix), *byh = dalloc(ix);
    double *prh = dalloc(ix), *vxh = dalloc(ix), *vyh = dalloc(ix), *ezh = dalloc(ix);
    double *dro = dalloc(ix), *dee = dalloc(ix), *drx = dalloc(ix), *dry = dalloc(ix), *dby = dalloc(ix);
    double *fx = dalloc(ix), *qx = dalloc(ix);

    for (int i = 1; i <= ix; i++) { /* :60-63 */
        dxim[I1(i)] = 1.0 / dxm[I1(i)];
        dxi[I1(i)] = 1.0 / dx[I1(i)];
    }
    for (int i = 1; i <= ix; i++) { /* :68-74 */
        dro[I1(i)] = 0.0; dee[I1(i)] = 0.0; drx[I1(i)] = 0.0; dry[I1(i)] = 0.0; dby[I1(i)] = 0.0;
    }
    for (int i = 1; i <= ix; i++) { /* :79-86, Bx not in the energy */
        double vv = sq(vx[I1(i)]) + sq(vy[I1(i)]);
        double bb = sq(by[I1(i)]);
        ee[I1(i)] = pr[I1(i)] / (gm - 1.0) + 0.5 * ro[I1(i)] * vv + pi8i * bb;
        rx[I1(i)] = ro[I1(i)] * vx[I1(i)];
        ry[I1(i)] = ro[I1(i)] * vy[I1(i)];
    }
    for (int i = 1; i <= ix; i++) /* :87-89 */
        ez[I1(i)] = -vx[I1(i)] * by[I1(i)] + vy[I1(i)] * bx[I1(i)];

    /* half step, :94-125 */
    for (int i = 1; i <= ix; i++) fx[I1(i)] = ro[I1(i)] * vx[I1(i)];
    mlwhalf1(ro, roh, dro, dt, fx, dxi, dxim, ix);
    for (int i = 1; i <= ix; i++) {
        double vv = sq(vx[I1(i)]) + sq(vy[I1(i)]);
        double ep = pr[I1(i)] * gm / (gm - 1.0) + 0.5 * ro[I1(i)] * vv;
        fx[I1(i)] = ep * vx[I1(i)] - by[I1(i)] * ez[I1(i)] * pi4i;
    }
    mlwhalf1(ee, eeh, dee, dt, fx, dxi, dxim, ix);
    for (int i = 1; i <= ix; i++) {
        double bbm = sq(by[I1(i)]);
        fx[I1(i)] = ro[I1(i)] * sq(vx[I1(i)]) + pr[I1(i)] + bbm * pi8i;
    }
    mlwhalf1(rx, rxh, drx, dt, fx, dxi, dxim, ix);
    for (int i = 1; i <= ix; i++)
        fx[I1(i)] = ro[I1(i)] * vx[I1(i)] * vy[I1(i)] - bx[I1(i)] * by[I1(i)] * pi4i;
    mlwhalf1(ry, ryh, dry, dt, fx, dxi, dxim, ix);
    for (int i = 1; i <= ix; i++) fx[I1(i)] = -ez[I1(i)];
    mlwhalf1(by, byh, dby, dt, fx, dxi, dxim, ix);

    /* half-step primitives, :129-139 */
    for (int i = 1; i <= ix - 1; i++) {
        vxh[I1(i)] = rxh[I1(i)] / roh[I1(i)];
        vyh[I1(i)] = ryh[I1(i)] / roh[I1(i)];
        double vv = sq(vxh[I1(i)]) + sq(vyh[I1(i)]);
        double bb = sq(byh[I1(i)]);
        prh[I1(i)] = (gm - 1.0) * (eeh[I1(i)] - 0.5 * roh[I1(i)] * vv - pi8i * bb);
    }
    for (int i = 1; i <= ix; i++)
        ezh[I1(i)] = -vxh[I1(i)] * byh[I1(i)] + vyh[I1(i)] * bxm[I1(i)];

    /* full step, :144-175 */
    for (int i = 1; i <= ix - 1; i++) fx[I1(i)] = roh[I1(i)] * vxh[I1(i)];
    mlwfull1(dro, dt, fx, dxi, ix);
    for (int i = 1; i <= ix - 1; i++) {
        double vv = sq(vxh[I1(i)]) + sq(vyh[I1(i)]);
        double eph = prh[I1(i)] * gm / (gm - 1.0) + 0.5 * roh[I1(i)] * vv;
        fx[I1(i)] = eph * vxh[I1(i)] - byh[I1(i)] * ezh[I1(i)] * pi4i;
    }
    mlwfull1(dee, dt, fx, dxi, ix);
    for (int i = 1; i <= ix - 1; i++) {
        double bbm = sq(byh[I1(i)]);
        fx[I1(i)] = roh[I1(i)] * sq(vxh[I1(i)]) + prh[I1(i)] + bbm * pi8i;
    }
    mlwfull1(drx, dt, fx, dxi, ix);
    for (int i = 1; i <= ix - 1; i++)
        fx[I1(i)] = roh[I1(i)] * vxh[I1(i)] * vyh[I1(i)] - bxm[I1(i)] * byh[I1(i)] * pi4i;
    mlwfull1(dry, dt, fx, dxi, ix);
    for (int i = 1; i <= ix - 1; i++) fx[I1(i)] = -ezh[I1(i)];
    mlwfull1(dby, dt, fx, dxi, ix);

    /* artificial viscosity, :180-191 */
    const double zero = 0.0;
    for (int i = 1; i <= ix - 1; i++)
        qx[I1(i)] = qav * dxm[I1(i)] * fmaxF(zero, fabs(vx[I1(i + 1)] - vx[I1(i)]) - R4_1EM4);
    mlwartv1(ro, dro, dt, qx, dxi, dxim, ix);
    mlwartv1(ee, dee, dt, qx, dxi, dxim, ix);
    mlwartv1(rx, drx, dt, qx, dxi, dxim, ix);
    mlwartv1(ry, dry, dt, qx, dxi, dxim, ix);
    mlwartv1(by, dby, dt, qx, dxi, dxim, ix);

    /* update and primitive recovery, :195-212 */
    for (int i = 2; i <= ix - 1; i++) {
        ro[I1(i)] = ro[I1(i)] + dro[I1(i)];
        rx[I1(i)] = rx[I1(i)] + drx[I1(i)];
        ry[I1(i)] = ry[I1(i)] + dry[I1(i)];
        by[I1(i)] = by[I1(i)] + dby[I1(i)];
        ee[I1(i)] = ee[I1(i)] + dee[I1(i)];
    }
    for (int i = 2; i <= ix - 1; i++) {
        vx[I1(i)] = rx[I1(i)] / ro[I1(i)];
        vy[I1(i)] = ry[I1(i)] / ro[I1(i)];
        double vv = sq(vx[I1(i)]) + sq(vy[I1(i)]);
        double bb = sq(by[I1(i)]);
        pr[I1(i)] = (gm - 1.0) * (ee[I1(i)] - 0.5 * ro[I1(i)] * vv - pi8i * bb);
    }
    free(dxi); free(dxim); free(ee); free(rx); free(ry); free(ez);
    free(roh); free(eeh); free(rxh); free(ryh); free(byh);
    free(prh); free(vxh); free(vyh); free(ezh);
    free(dro); free(dee); free(drx); free(dry); free(dby); free(fx); free(qx);
}

/* common/cfl_m.f:42-71 (interior only, |vx| + fast speed) */
void or1d_cfl_m(double *dt, double safety, double dtmin, int *merr, double gm,
                const double *bx, const double *ro, const double *pr,
                const double *vx, const double *vy, const double *by,
                const double *dx, int ix, int *imin) {
    (void)vy;
    const double pi = acos(-1.0);
    const double pi4i = 0.25 / pi;
    double *dtq = dalloc(ix);
    double d = R4_1E20;
    int im = 0;
    for (int i = 2; i <= ix - 1; i++) {
        double onero = 1.0 / ro[I1(i)];
        double ca2 = (bx[I1(i)] * bx[I1(i)] + by[I1(i)] * by[I1(i)]) * pi4i * onero;
        double cs2 = gm * pr[I1(i)] * onero;
        double dtcfl = dx[I1(i)] / (fabs(vx[I1(i)]) + sqrt(cs2 + ca2));
        dtq[I1(i)] = safety * dtcfl;
    }
    for (int i = 2; i <= ix - 1; i++)
        if (dtq[I1(i)] < d) { im = i; d = dtq[I1(i)]; }
    *merr = (d < dtmin) ? 9001 : 0;
    *dt = d;
    if (imin) *imin = im;
    free(dtq);
}

/* bc/bd{fre,spp,spn,smp,smn,cns,frd}x.f (side mbnd: 0 low, 1 high) */
void or1d_bd(int kind, int mbnd, int margin, double *qq, double q0,
             const double *dxm, int ix) {
    if (mbnd == 0) {
        int ibnd = 1 + margin;
        double dqq = 0.0;
        if (kind == 6) dqq = (qq[I1(ibnd + 1)] - qq[I1(ibnd)]) / dxm[I1(ibnd)]; /* bdfrdx.f */
        for (int i = 1; i <= margin; i++) {
            switch (kind) {
            case 0: qq[I1(ibnd - i)] = qq[I1(ibnd)]; break;          /* bdfrex.f:30-33 */
            case 1: qq[I1(ibnd - i)] = qq[I1(ibnd - 1 + i)]; break;  /* bdsppx.f */
            case 2: qq[I1(ibnd - i)] = -qq[I1(ibnd - 1 + i)]; break; /* bdspnx.f */
            case 3: qq[I1(ibnd - i)] = qq[I1(ibnd + i)]; break;      /* bdsmpx.f */
            case 4: qq[I1(ibnd - i)] = -qq[I1(ibnd + i)]; break;     /* bdsmnx.f:33-38 */
            case 5: qq[I1(ibnd - i)] = q0; break;                    /* bdcnsx.f */
            case 6: qq[I1(ibnd - i)] = qq[I1(ibnd - i + 1)] - dqq * dxm[I1(ibnd - i)]; break;
            }
        }
        if (kind == 4) qq[I1(ibnd)] = 0.0; /* bdsmnx.f: wall value */
    } else {
        int ibnd = ix - margin;
        double dqq = 0.0;
        if (kind == 6) dqq = (qq[I1(ibnd)] - qq[I1(ibnd - 1)]) / dxm[I1(ibnd - 1)];
        for (int i = 1; i <= margin; i++) {
            switch (kind) {
            case 0: qq[I1(ibnd + i)] = qq[I1(ibnd)]; break;
            case 1: qq[I1(ibnd + i)] = qq[I1(ibnd + 1 - i)]; break;
            case 2: qq[I1(ibnd + i)] = -qq[I1(ibnd + 1 - i)]; break;
            case 3: qq[I1(ibnd + i)] = qq[I1(ibnd - i)]; break;
            case 4: qq[I1(ibnd + i)] = -qq[I1(ibnd - i)]; break;
            case 5: qq[I1(ibnd + i)] = q0; break;
            case 6: qq[I1(ibnd + i)] = qq[I1(ibnd + i - 1)] + dqq * dxm[I1(ibnd + i - 1)]; break;
            }
        }
        if (kind == 4) qq[I1(ibnd)] = 0.0;
    }
}

/* bc/bdperx.f:31-40 */
void or1d_bdperx(int margin0, int margin1, double *qq, int ix) {
    int ibnd0 = 1 + margin0, ibnd1 = ix - margin1;
    for (int i = 1; i <= margin0; i++) qq[I1(ibnd0 - i)] = qq[I1(ibnd1 + 1 - i)];
    for (int i = 1; i <= margin1; i++) qq[I1(ibnd1 + i)] = qq[I1(ibnd0 - 1 + i)];
}

/* common/grdrdy.f:28-42 */
void or1d_grdrdy(double *dx, double *xm, double *dxm, const double *x, int ix) {
    for (int i = 1; i <= ix - 1; i++) dxm[I1(i)] = x[I1(i + 1)] - x[I1(i)];
    dxm[I1(ix)] = dxm[I1(ix - 1)];
    for (int i = 2; i <= ix - 1; i++) dx[I1(i)] = (dxm[I1(i - 1)] + dxm[I1(i)]) / 2.0;
    dx[I1(1)] = dx[I1(2)];
    dx[I1(ix)] = dx[I1(ix - 1)];
    for (int i = 1; i <= ix - 1; i++) xm[I1(i)] = (x[I1(i)] + x[I1(i + 1)]) / 2.0;
    xm[I1(ix)] = xm[I1(ix - 1)] + dx[I1(ix - 1)];
}

/* hdmlw/mlwsrch.f:32-41: source at the cell (du) and at the half step (un) */
static void mlwsrch1(double *un, double *du, double dt, const double *s, int ix) {
    for (int i = 2; i <= ix - 1; i++) du[I1(i)] = du[I1(i)] + 0.5 * dt * s[I1(i)];
    for (int i = 1; i <= ix - 1; i++) {
        double sh = 0.5 * (s[I1(i + 1)] + s[I1(i)]);
        un[I1(i)] = un[I1(i)] + dt * sh;
    }
}

/* hdmlw/mlwsrcf.f:31-34: half-step source interpolated to the cell */
static void mlwsrcf1(double *du, double dt, const double *s, const double *ux0,
                     const double *ux1, int ix) {
    for (int i = 2; i <= ix - 1; i++) {
        double sh = ux0[I1(i)] * s[I1(i - 1)] + ux1[I1(i)] * s[I1(i)];
        du[I1(i)] = du[I1(i)] + 0.5 * dt * sh;
    }
}

/* hdmlw/mlw_m_bg.f:2-271: mlw_m with gravity, flux-tube cross-section sc and
 * rotation radius rr (conserved variables carried as u*sc, ry*sc*rr,
 * by*sc/rr); Bx in the energy. */
void or1d_mlw_m_bg(double *ro, double *pr, double *vx, double *vy, double *by,
                   const double *bx, const double *bxm, double dt, double qav,
                   double gm, const double *gx, const double *gxm,
                   const double *sc, const double *dsc, const double *scm,
                   const double *dscm, const double *rr, const double *rrm,
                   const double *drr, const double *drrm, const double *dx,
                   const double *dxm, int ix) {
    const double pi = acos(-1.0);
    const double pi8i = 1.0 / pi / 8.0; /* :70 */
    const double pi4i = 1.0 / pi / 4.0; /* :71 */
    double *dxi = dalloc(ix), *dxim = dalloc(ix), *ux0 = dalloc(ix), *ux1 = dalloc(ix);
    double *ee = dalloc(ix), *rx = dalloc(ix), *ry = dalloc(ix), *ez = dalloc(ix);
    double *roh = dalloc(ix), *prh = dalloc(ix), *vxh = dalloc(ix), *vyh = dalloc(ix), *byh = dalloc(ix);
    double *eeh = dalloc(ix), *ezh = dalloc(ix);
    double *rosc = dalloc(ix), *eesc = dalloc(ix), *rxsc = dalloc(ix), *rysc = dalloc(ix), *bysc = dalloc(ix);
    double *rosch = dalloc(ix), *eesch = dalloc(ix), *rxsch = dalloc(ix), *rysch = dalloc(ix), *bysch = dalloc(ix);
    double *drosc = dalloc(ix), *deesc = dalloc(ix), *drxsc = dalloc(ix), *drysc = dalloc(ix), *dbysc = dalloc(ix);
    double *fx = dalloc(ix), *qx = dalloc(ix), *ss = dalloc(ix);

    for (int i = 1; i <= ix; i++) { /* :75-78 */
        dxim[I1(i)] = 1.0 / dxm[I1(i)];
        dxi[I1(i)] = 1.0 / dx[I1(i)];
    }
    for (int i = 2; i <= ix - 1; i++) { /* :79-82 */
        ux1[I1(i)] = 0.5 * dxm[I1(i - 1)] / dx[I1(i)];
        ux0[I1(i)] = 0.5 * dxm[I1(i)] / dx[I1(i)];
    }
    for (int i = 1; i <= ix; i++) { /* :87-93 */
        drosc[I1(i)] = 0.0; deesc[I1(i)] = 0.0; drxsc[I1(i)] = 0.0; drysc[I1(i)] = 0.0; dbysc[I1(i)] = 0.0;
    }
    for (int i = 1; i <= ix; i++) { /* :97-104, Bx in the energy */
        double vv = sq(vx[I1(i)]) + sq(vy[I1(i)]);
        double bb = sq(by[I1(i)]) + sq(bx[I1(i)]);
        ee[I1(i)] = pr[I1(i)] / (gm - 1.0) + 0.5 * ro[I1(i)] * vv + pi8i * bb;
        rx[I1(i)] = ro[I1(i)] * vx[I1(i)];
        ry[I1(i)] = ro[I1(i)] * vy[I1(i)];
    }
    for (int i = 1; i <= ix; i++) ez[I1(i)] = -vx[I1(i)] * by[I1(i)] + vy[I1(i)] * bx[I1(i)]; /* :105-107 */
    for (int i = 1; i <= ix; i++) { /* :108-114 */
        rosc[I1(i)] = ro[I1(i)] * sc[I1(i)];
        eesc[I1(i)] = ee[I1(i)] * sc[I1(i)];
        rxsc[I1(i)] = rx[I1(i)] * sc[I1(i)];
        rysc[I1(i)] = ry[I1(i)] * sc[I1(i)] * rr[I1(i)];
        bysc[I1(i)] = by[I1(i)] * sc[I1(i)] / rr[I1(i)];
    }

    /* half step, :120-167 */
    for (int i = 1; i <= ix; i++) {
        fx[I1(i)] = ro[I1(i)] * vx[I1(i)];
        fx[I1(i)] = fx[I1(i)] * sc[I1(i)];
    }
    mlwhalf1(rosc, rosch, drosc, dt, fx, dxi, dxim, ix);
    for (int i = 1; i <= ix; i++) {
        double vv = sq(vx[I1(i)]) + sq(vy[I1(i)]);
        double ep = pr[I1(i)] * gm / (gm - 1.0) + 0.5 * ro[I1(i)] * vv;
        fx[I1(i)] = ep * vx[I1(i)] - by[I1(i)] * ez[I1(i)] * pi4i;
        fx[I1(i)] = fx[I1(i)] * sc[I1(i)];
        ss[I1(i)] = ro[I1(i)] * vx[I1(i)] * gx[I1(i)];
        ss[I1(i)] = ss[I1(i)] * sc[I1(i)];
    }
    mlwhalf1(eesc, eesch, deesc, dt, fx, dxi, dxim, ix);
    mlwsrch1(eesch, deesc, dt, ss, ix);
    for (int i = 1; i <= ix; i++) {
        double bbm = sq(by[I1(i)]) - sq(bx[I1(i)]);
        double bb = sq(by[I1(i)]) + sq(bx[I1(i)]);
        fx[I1(i)] = ro[I1(i)] * sq(vx[I1(i)]) + pr[I1(i)] + bbm * pi8i;
        ss[I1(i)] = ro[I1(i)] * gx[I1(i)]
                  + (ro[I1(i)] * sq(vy[I1(i)]) - sq(by[I1(i)]) * pi4i) / rr[I1(i)] * drr[I1(i)];
        fx[I1(i)] = fx[I1(i)] * sc[I1(i)];
        ss[I1(i)] = ss[I1(i)] * sc[I1(i)] + (pr[I1(i)] + bb * pi8i) * dsc[I1(i)];
    }
    mlwhalf1(rxsc, rxsch, drxsc, dt, fx, dxi, dxim, ix);
    mlwsrch1(rxsch, drxsc, dt, ss, ix);
    for (int i = 1; i <= ix; i++) {
        fx[I1(i)] = ro[I1(i)] * vx[I1(i)] * vy[I1(i)] - bx[I1(i)] * by[I1(i)] * pi4i;
        fx[I1(i)] = fx[I1(i)] * sc[I1(i)] * rr[I1(i)];
    }
    mlwhalf1(rysc, rysch, drysc, dt, fx, dxi, dxim, ix);
    for (int i = 1; i <= ix; i++) {
        fx[I1(i)] = -ez[I1(i)];
        fx[I1(i)] = fx[I1(i)] * sc[I1(i)];
        fx[I1(i)] = fx[I1(i)] / rr[I1(i)];
    }
    mlwhalf1(bysc, bysch, dbysc, dt, fx, dxi, dxim, ix);

    /* half-step primitives, :170-183 */
    for (int i = 1; i <= ix - 1; i++) {
        roh[I1(i)] = rosch[I1(i)] / scm[I1(i)];
        eeh[I1(i)] = eesch[I1(i)] / scm[I1(i)];
        vxh[I1(i)] = rxsch[I1(i)] / rosch[I1(i)];
        vyh[I1(i)] = rysch[I1(i)] / rosch[I1(i)] / rrm[I1(i)];
        byh[I1(i)] = bysch[I1(i)] / scm[I1(i)] * rrm[I1(i)];
        double vv = sq(vxh[I1(i)]) + sq(vyh[I1(i)]);
        double bb = sq(bxm[I1(i)]) + sq(byh[I1(i)]);
        prh[I1(i)] = (gm - 1.0) * (eeh[I1(i)] - 0.5 * roh[I1(i)] * vv - pi8i * bb);
    }
    /* :181-183 loops to ix; element ix reads the never-written vxh(ix) etc.
     * and is never used, so the restatement stops at ix-1 */
    for (int i = 1; i <= ix - 1; i++) ezh[I1(i)] = -vxh[I1(i)] * byh[I1(i)] + vyh[I1(i)] * bxm[I1(i)];

    /* full step, :188-234 */
    for (int i = 1; i <= ix - 1; i++) {
        fx[I1(i)] = roh[I1(i)] * vxh[I1(i)];
        fx[I1(i)] = fx[I1(i)] * scm[I1(i)];
    }
    mlwfull1(drosc, dt, fx, dxi, ix);
    for (int i = 1; i <= ix - 1; i++) {
        double vv = sq(vxh[I1(i)]) + sq(vyh[I1(i)]);
        double eph = prh[I1(i)] * gm / (gm - 1.0) + 0.5 * roh[I1(i)] * vv;
        fx[I1(i)] = eph * vxh[I1(i)] - byh[I1(i)] * ezh[I1(i)] * pi4i;
        ss[I1(i)] = roh[I1(i)] * vxh[I1(i)] * gxm[I1(i)];
        fx[I1(i)] = fx[I1(i)] * scm[I1(i)];
        ss[I1(i)] = ss[I1(i)] * scm[I1(i)];
    }
    mlwfull1(deesc, dt, fx, dxi, ix);
    mlwsrcf1(deesc, dt, ss, ux0, ux1, ix);
    for (int i = 1; i <= ix - 1; i++) {
        double bbm = sq(byh[I1(i)]) - sq(bxm[I1(i)]);
        double bb = sq(byh[I1(i)]) + sq(bxm[I1(i)]);
        fx[I1(i)] = roh[I1(i)] * sq(vxh[I1(i)]) + prh[I1(i)] + bbm * pi8i;
        ss[I1(i)] = roh[I1(i)] * gxm[I1(i)]
                  + (roh[I1(i)] * sq(vyh[I1(i)]) - sq(byh[I1(i)]) * pi4i) / rrm[I1(i)] * drrm[I1(i)];
        fx[I1(i)] = fx[I1(i)] * scm[I1(i)];
        ss[I1(i)] = ss[I1(i)] * scm[I1(i)] + (prh[I1(i)] + bb * pi8i) * dscm[I1(i)];
    }
    mlwfull1(drxsc, dt, fx, dxi, ix);
    mlwsrcf1(drxsc, dt, ss, ux0, ux1, ix);
    for (int i = 1; i <= ix - 1; i++) {
        fx[I1(i)] = roh[I1(i)] * vxh[I1(i)] * vyh[I1(i)] - bxm[I1(i)] * byh[I1(i)] * pi4i;
        fx[I1(i)] = fx[I1(i)] * scm[I1(i)] * rrm[I1(i)];
    }
    mlwfull1(drysc, dt, fx, dxi, ix);
    for (int i = 1; i <= ix - 1; i++) {
        fx[I1(i)] = -ezh[I1(i)];
        fx[I1(i)] = fx[I1(i)] * scm[I1(i)];
        fx[I1(i)] = fx[I1(i)] / rrm[I1(i)];
    }
    mlwfull1(dbysc, dt, fx, dxi, ix);

    /* artificial viscosity on the sc-weighted variables, :240-247 */
    const double zero = 0.0;
    for (int i = 1; i <= ix - 1; i++)
        qx[I1(i)] = qav * dxm[I1(i)] * fmaxF(zero, fabs(vx[I1(i + 1)] - vx[I1(i)]) - R4_1EM4);
    mlwartv1(rosc, drosc, dt, qx, dxi, dxim, ix);
    mlwartv1(eesc, deesc, dt, qx, dxi, dxim, ix);
    mlwartv1(rxsc, drxsc, dt, qx, dxi, dxim, ix);
    mlwartv1(rysc, drysc, dt, qx, dxi, dxim, ix);
    mlwartv1(bysc, dbysc, dt, qx, dxi, dxim, ix);

    /* update and primitive recovery, :251-268 */
    for (int i = 2; i <= ix - 1; i++) {
        ro[I1(i)] = (rosc[I1(i)] + drosc[I1(i)]) / sc[I1(i)];
        ee[I1(i)] = (eesc[I1(i)] + deesc[I1(i)]) / sc[I1(i)];
        rx[I1(i)] = (rxsc[I1(i)] + drxsc[I1(i)]) / sc[I1(i)];
        ry[I1(i)] = (rysc[I1(i)] + drysc[I1(i)]) / sc[I1(i)] / rr[I1(i)];
        by[I1(i)] = (bysc[I1(i)] + dbysc[I1(i)]) / sc[I1(i)] * rr[I1(i)];
    }
    for (int i = 2; i <= ix - 1; i++) {
        vx[I1(i)] = rx[I1(i)] / ro[I1(i)];
        vy[I1(i)] = ry[I1(i)] / ro[I1(i)];
        double vv = sq(vx[I1(i)]) + sq(vy[I1(i)]);
        double bb = sq(bx[I1(i)]) + sq(by[I1(i)]);
        pr[I1(i)] = (gm - 1.0) * (ee[I1(i)] - 0.5 * ro[I1(i)] * vv - pi8i * bb);
    }
    free(dxi); free(dxim); free(ux0); free(ux1); free(ee); free(rx); free(ry); free(ez);
    free(roh); free(prh); free(vxh); free(vyh); free(byh); free(eeh); free(ezh);
    free(rosc); free(eesc); free(rxsc); free(rysc); free(bysc);
    free(rosch); free(eesch); free(rxsch); free(rysch); free(bysch);
    free(drosc); free(deesc); free(drxsc); free(drysc); free(dbysc);
    free(fx); free(qx); free(ss);
}

/* hdmlw/mlw_m3.f:2-252: 1D MHD with three velocity and field components
 * (Bx constant, not in the energy) */
void or1d_mlw_m3(double *ro, double *pr, double *vx, double *vy, double *vz,
                 double *by, double *bz, const double *bx, const double *bxm,
                 double dt, double qav, double gm, const double *dx,
                 const double *dxm, int ix) {
    const double pi = acos(-1.0);
    const double pi8i = 1.0 / pi / 8.0; /* :57 */
    const double pi4i = 1.0 / pi / 4.0;
    double *dxi = dalloc(ix), *dxim = dalloc(ix);
    double *ee = dalloc(ix), *rx = dalloc(ix), *ry = dalloc(ix), *rz = dalloc(ix), *ey = dalloc(ix), *ez = dalloc(ix);
    double *roh = dalloc(ix), *eeh = dalloc(ix), *rxh = dalloc(ix), *ryh = dalloc(ix), *rzh = dalloc(ix);
    double *byh = dalloc(ix), *bzh = dalloc(ix), *prh = dalloc(ix), *vxh = dalloc(ix), *vyh = dalloc(ix);
    double *vzh = dalloc(ix), *eyh = dalloc(ix), *ezh = dalloc(ix);
    double *dro = dalloc(ix), *dee = dalloc(ix), *drx = dalloc(ix), *dry = dalloc(ix), *drz = dalloc(ix);
    double *dby = dalloc(ix), *dbz = dalloc(ix), *fx = dalloc(ix), *qx = dalloc(ix);
    for (int i = 1; i <= ix; i++) {
        dxim[I1(i)] = 1.0 / dxm[I1(i)];
        dxi[I1(i)] = 1.0 / dx[I1(i)];
    }
    for (int i = 1; i <= ix; i++) { /* :71-79 */
        dro[I1(i)] = 0.0; dee[I1(i)] = 0.0; drx[I1(i)] = 0.0; dry[I1(i)] = 0.0;
        dby[I1(i)] = 0.0; drz[I1(i)] = 0.0; dbz[I1(i)] = 0.0;
    }
    for (int i = 1; i <= ix; i++) { /* :83-91 */
        double vv = sq(vx[I1(i)]) + sq(vy[I1(i)]) + sq(vz[I1(i)]);
        double bb = sq(by[I1(i)]) + sq(bz[I1(i)]);
        ee[I1(i)] = pr[I1(i)] / (gm - 1.0) + 0.5 * ro[I1(i)] * vv + pi8i * bb;
        rx[I1(i)] = ro[I1(i)] * vx[I1(i)];
        ry[I1(i)] = ro[I1(i)] * vy[I1(i)];
        rz[I1(i)] = ro[I1(i)] * vz[I1(i)];
    }
    for (int i = 1; i <= ix; i++) { /* :92-95 */
        ey[I1(i)] = -vz[I1(i)] * bx[I1(i)] + vx[I1(i)] * bz[I1(i)];
        ez[I1(i)] = -vx[I1(i)] * by[I1(i)] + vy[I1(i)] * bx[I1(i)];
    }
    /* half step, :99-143 */
    for (int i = 1; i <= ix; i++) fx[I1(i)] = ro[I1(i)] * vx[I1(i)];
    mlwhalf1(ro, roh, dro, dt, fx, dxi, dxim, ix);
    for (int i = 1; i <= ix; i++) {
        double vv = sq(vx[I1(i)]) + sq(vy[I1(i)]) + sq(vz[I1(i)]);
        double ep = pr[I1(i)] * gm / (gm - 1.0) + 0.5 * ro[I1(i)] * vv;
        fx[I1(i)] = ep * vx[I1(i)] - by[I1(i)] * ez[I1(i)] * pi4i + bz[I1(i)] * ey[I1(i)] * pi4i;
    }
    mlwhalf1(ee, eeh, dee, dt, fx, dxi, dxim, ix);
    for (int i = 1; i <= ix; i++) {
        double bbm = sq(by[I1(i)]) + sq(bz[I1(i)]);
        fx[I1(i)] = ro[I1(i)] * sq(vx[I1(i)]) + pr[I1(i)] + bbm * pi8i;
    }
    mlwhalf1(rx, rxh, drx, dt, fx, dxi, dxim, ix);
    for (int i = 1; i <= ix; i++) fx[I1(i)] = ro[I1(i)] * vx[I1(i)] * vy[I1(i)] - bx[I1(i)] * by[I1(i)] * pi4i;
    mlwhalf1(ry, ryh, dry, dt, fx, dxi, dxim, ix);
    for (int i = 1; i <= ix; i++) fx[I1(i)] = ro[I1(i)] * vx[I1(i)] * vz[I1(i)] - bx[I1(i)] * bz[I1(i)] * pi4i;
    mlwhalf1(rz, rzh, drz, dt, fx, dxi, dxim, ix);
    for (int i = 1; i <= ix; i++) fx[I1(i)] = -ez[I1(i)];
    mlwhalf1(by, byh, dby, dt, fx, dxi, dxim, ix);
    for (int i = 1; i <= ix; i++) fx[I1(i)] = ey[I1(i)];
    mlwhalf1(bz, bzh, dbz, dt, fx, dxi, dxim, ix);
    /* half-step primitives, :147-159 */
    for (int i = 1; i <= ix - 1; i++) {
        vxh[I1(i)] = rxh[I1(i)] / roh[I1(i)];
        vyh[I1(i)] = ryh[I1(i)] / roh[I1(i)];
        vzh[I1(i)] = rzh[I1(i)] / roh[I1(i)];
        double vv = sq(vxh[I1(i)]) + sq(vyh[I1(i)]) + sq(vzh[I1(i)]);
        double bb = sq(byh[I1(i)]) + sq(bzh[I1(i)]);
        prh[I1(i)] = (gm - 1.0) * (eeh[I1(i)] - 0.5 * roh[I1(i)] * vv - pi8i * bb);
    }
    for (int i = 1; i <= ix - 1; i++) {
        eyh[I1(i)] = -vzh[I1(i)] * bxm[I1(i)] + vxh[I1(i)] * bzh[I1(i)];
        ezh[I1(i)] = -vxh[I1(i)] * byh[I1(i)] + vyh[I1(i)] * bxm[I1(i)];
    }
    /* full step, :163-207 */
    for (int i = 1; i <= ix - 1; i++) fx[I1(i)] = roh[I1(i)] * vxh[I1(i)];
    mlwfull1(dro, dt, fx, dxi, ix);
    for (int i = 1; i <= ix - 1; i++) {
        double vv = sq(vxh[I1(i)]) + sq(vyh[I1(i)]) + sq(vzh[I1(i)]);
        double eph = prh[I1(i)] * gm / (gm - 1.0) + 0.5 * roh[I1(i)] * vv;
        fx[I1(i)] = eph * vxh[I1(i)] - byh[I1(i)] * ezh[I1(i)] * pi4i + bzh[I1(i)] * eyh[I1(i)] * pi4i;
    }
    mlwfull1(dee, dt, fx, dxi, ix);
    for (int i = 1; i <= ix - 1; i++) {
        double bbm = sq(byh[I1(i)]) + sq(bzh[I1(i)]);
        fx[I1(i)] = roh[I1(i)] * sq(vxh[I1(i)]) + prh[I1(i)] + bbm * pi8i;
    }
    mlwfull1(drx, dt, fx, dxi, ix);
    for (int i = 1; i <= ix - 1; i++)
        fx[I1(i)] = roh[I1(i)] * vxh[I1(i)] * vyh[I1(i)] - bxm[I1(i)] * byh[I1(i)] * pi4i;
    mlwfull1(dry, dt, fx, dxi, ix);
    for (int i = 1; i <= ix - 1; i++)
        fx[I1(i)] = roh[I1(i)] * vxh[I1(i)] * vzh[I1(i)] - bxm[I1(i)] * bzh[I1(i)] * pi4i;
    mlwfull1(drz, dt, fx, dxi, ix);
    for (int i = 1; i <= ix - 1; i++) fx[I1(i)] = -ezh[I1(i)];
    mlwfull1(dby, dt, fx, dxi, ix);
    for (int i = 1; i <= ix - 1; i++) fx[I1(i)] = eyh[I1(i)];
    mlwfull1(dbz, dt, fx, dxi, ix);
    /* artificial viscosity, :212-225 */
    const double zero = 0.0;
    for (int i = 1; i <= ix - 1; i++)
        qx[I1(i)] = qav * dxm[I1(i)] * fmaxF(zero, fabs(vx[I1(i + 1)] - vx[I1(i)]) - R4_1EM4);
    mlwartv1(ro, dro, dt, qx, dxi, dxim, ix);
    mlwartv1(ee, dee, dt, qx, dxi, dxim, ix);
    mlwartv1(rx, drx, dt, qx, dxi, dxim, ix);
    mlwartv1(ry, dry, dt, qx, dxi, dxim, ix);
    mlwartv1(rz, drz, dt, qx, dxi, dxim, ix);
    mlwartv1(by, dby, dt, qx, dxi, dxim, ix);
    mlwartv1(bz, dbz, dt, qx, dxi, dxim, ix);
    /* update and primitive recovery, :229-249 */
    for (int i = 2; i <= ix - 1; i++) {
        ro[I1(i)] = ro[I1(i)] + dro[I1(i)];
        rx[I1(i)] = rx[I1(i)] + drx[I1(i)];
        ry[I1(i)] = ry[I1(i)] + dry[I1(i)];
        rz[I1(i)] = rz[I1(i)] + drz[I1(i)];
        by[I1(i)] = by[I1(i)] + dby[I1(i)];
        bz[I1(i)] = bz[I1(i)] + dbz[I1(i)];
        ee[I1(i)] = ee[I1(i)] + dee[I1(i)];
    }
    for (int i = 2; i <= ix - 1; i++) {
        vx[I1(i)] = rx[I1(i)] / ro[I1(i)];
        vy[I1(i)] = ry[I1(i)] / ro[I1(i)];
        vz[I1(i)] = rz[I1(i)] / ro[I1(i)];
        double vv = sq(vx[I1(i)]) + sq(vy[I1(i)]) + sq(vz[I1(i)]);
        double bb = sq(by[I1(i)]) + sq(bz[I1(i)]);
        pr[I1(i)] = (gm - 1.0) * (ee[I1(i)] - 0.5 * ro[I1(i)] * vv - pi8i * bb);
    }
    free(dxi); free(dxim); free(ee); free(rx); free(ry); free(rz); free(ey); free(ez);
    free(roh); free(eeh); free(rxh); free(ryh); free(rzh); free(byh); free(bzh); free(prh);
    free(vxh); free(vyh); free(vzh); free(eyh); free(ezh);
    free(dro); free(dee); free(drx); free(dry); free(drz); free(dby); free(dbz); free(fx); free(qx);
}

/* common/cfl_m3.f:42-71 (1D, interior only, |vx| + fast speed with bz) */
void or1d_cfl_m3(double *dt, double safety, double dtmin, int *merr, double gm,
                 const double *bx, const double *ro, const double *pr,
                 const double *vx, const double *vy, const double *by,
                 const double *vz, const double *bz, const double *dx, int ix,
                 int *imin) {
    (void)vy; (void)vz;
    const double pi = acos(-1.0);
    const double pi4i = 0.25 / pi;
    double *dtq = dalloc(ix);
    double d = R4_1E20;
    int im = 0;
    for (int i = 2; i <= ix - 1; i++) {
        double onero = 1.0 / ro[I1(i)];
        double b2 = sq(bx[I1(i)]) + sq(by[I1(i)]) + sq(bz[I1(i)]);
        double ca2 = b2 * pi4i * onero;
        double cs2 = gm * pr[I1(i)] * onero;
        double dtcfl = dx[I1(i)] / (fabs(vx[I1(i)]) + sqrt(cs2 + ca2));
        dtq[I1(i)] = safety * dtcfl;
    }
    for (int i = 2; i <= ix - 1; i++)
        if (dtq[I1(i)] < d) { im = i; d = dtq[I1(i)]; }
    *merr = (d < dtmin) ? 9001 : 0;
    *dt = d;
    if (imin) *imin = im;
    free(dtq);
}

/* ---- the helper routines themselves, with the reference argument lists
 * (the names the drop-in library exports; see cans_oracle.h) */
void or1d_mlwhalf(const double *u, double *un, double *du, double dt, const double *f, const double *dxi,
                  const double *dxim, int ix) {
    mlwhalf1(u, un, du, dt, f, dxi, dxim, ix);
}
void or1d_mlwfull(double *du, double dt, const double *f, const double *dxi, int ix) {
    mlwfull1(du, dt, f, dxi, ix);
}
void or1d_mlwartv(const double *u, double *du, double dt, const double *qx, const double *dxi, const double *dxim,
                  int ix) {
    mlwartv1(u, du, dt, qx, dxi, dxim, ix);
}
void or1d_mlwsrch(double *un, double *du, double dt, const double *s, int ix) { mlwsrch1(un, du, dt, s, ix); }
void or1d_mlwsrcf(double *du, double dt, const double *s, const double *ux0, const double *ux1, int ix) {
    mlwsrcf1(du, dt, s, ux0, ux1, ix);
}
