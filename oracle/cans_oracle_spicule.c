/*
 * cans_oracle_spicule.c -- C restatement of the UTRIP spicule time loop
 * around mlw_m_bg (test infrastructure only; see cans_oracle.h).
 * Citations are to cans_beta_vary/cans1d/.
 *
 * The reference's torque source draws its noise from the compiler's
 * RANDOM_NUMBER, whose stream is not reproducible across compilers.  The
 * restatement draws it from the Park-Miller generator of common/rangen.f
 * instead (one draw per cell per call, cell order), with the seed carried by
 * the caller.
 */
#include "cans_oracle.h"
#include "cans_oracle_util.h"

/* common/scrdy.f:28-34 (elements 1, ix of dsc and 1, ix-1, ix of dscm are
 * left as they are) */
void or1d_scrdy(double *dsc, double *dscm, const double *sc, const double *scm,
                const double *dx, const double *dxm, int ix) {
    for (int i = 2; i <= ix - 1; i++) dsc[I1(i)] = (scm[I1(i)] - scm[I1(i - 1)]) / dx[I1(i)];
    for (int i = 2; i <= ix - 2; i++) dscm[I1(i)] = (sc[I1(i + 1)] - sc[I1(i)]) / dxm[I1(i)];
}

/* common/chkdav.f:31-39 */
void or1d_chkdav(int *n_events, double *da, double *vx, double floor_, int ix) {
    int n = 0;
    for (int i = 1; i <= ix; i++) {
        if (da[I1(i)] < floor_) {
            n = n + 1;
            da[I1(i)] = floor_;
            vx[I1(i)] = 0.0;
        }
    }
    *n_events = n;
}

/* common/rangen.f:11-21: Park-Miller minimal standard by Schrage's method */
double or_rangen(int *mseed) {
    const int ma = 16807, mm = 2147483647, mq = 127773, mr = 2836;
    const double rnml = 4.656612875245796924105750827168e-10;
    int k = *mseed / mq;
    int s = ma * (*mseed - k * mq) - mr * k;
    if (s < 0) s = s + mm;
    *mseed = s;
    return rnml * (double)s;
}

/* md_spicule/hdmlw/trqalfven.f:23-38 with RANDOM_NUMBER replaced by rangen */
void orsp_trqalfven(double *vy, double dt, double time, const double *x, int ix,
                    const double *rr, int *mseed) {
    (void)time;
    double *q = dalloc(ix), *trq = dalloc(ix);
    for (int i = 1; i <= ix; i++) {
        double anum = or_rangen(mseed);
        q[I1(i)] = 2.0 * rr[I1(i)] * (anum - 0.5);
        trq[I1(i)] = tanh((x[I1(i)] - 0.75) / (double)0.075f) - 1.0;
    }
    trq[I1(ix)] = trq[I1(ix - 2)];
    q[I1(ix)] = q[I1(ix - 2)];
    const double trq0 = 30.0;
    for (int i = 1; i <= ix; i++) vy[I1(i)] = vy[I1(i)] + dt * trq0 * q[I1(i)] * trq[I1(i)];
    free(q);
    free(trq);
}

/* md_spicule/hdmlw/bnd.f:10-20 */
void orsp_bnd(int margin, double *ro, double *pr, double *vx, double *vy,
              double *by, int ix) {
    or1d_bd(1, 0, margin, pr, 0.0, NULL, ix);
    or1d_bd(1, 0, margin, ro, 0.0, NULL, ix);
    or1d_bd(2, 0, margin, vx, 0.0, NULL, ix);
    or1d_bd(1, 0, margin, vy, 0.0, NULL, ix);
    or1d_bd(1, 0, margin, by, 0.0, NULL, ix);
    or1d_bd(0, 1, margin, pr, 0.0, NULL, ix);
    or1d_bd(0, 1, margin, ro, 0.0, NULL, ix);
    or1d_bd(0, 1, margin, vx, 0.0, NULL, ix);
    or1d_bd(0, 1, margin, vy, 0.0, NULL, ix);
    or1d_bd(0, 1, margin, by, 0.0, NULL, ix);
}

/* md_spicule/hdmlw/bndsc.f:10-18 */
void orsp_bndsc(int margin, double *sc, double *dsc, double *scm, double *dscm,
                int ix) {
    or1d_bd(1, 0, margin, sc, 0.0, NULL, ix);
    or1d_bd(2, 0, margin, dsc, 0.0, NULL, ix);
    or1d_bd(3, 0, margin - 1, scm, 0.0, NULL, ix);
    or1d_bd(4, 0, margin - 1, dscm, 0.0, NULL, ix);
    or1d_bd(1, 1, margin, sc, 0.0, NULL, ix);
    or1d_bd(2, 1, margin, dsc, 0.0, NULL, ix);
    or1d_bd(3, 1, margin, scm, 0.0, NULL, ix);
    or1d_bd(4, 1, margin, dscm, 0.0, NULL, ix);
}

/* md_spicule/hdmlw/main.f:169-227, the time loop without its output:
 *   ns=ns+1; cfl_m (merr -> stop); tp=t; t=t+dt; mlw_m_bg; trqalfven; bnd;
 *   chkdav(ro); chkdav(pr); loop while ns < nstop and t < tend.
 * safety 0.8, dtmin 1e-10, qav 1, floor 1e-9 are main.f's; they are
 * arguments here.  The geometry g[] holds bx, bxm, gx, gxm, sc, dsc, scm,
 * dscm, rr, rrm, drr, drrm, dx, dxm, x (in that order). */
void orsp_run(double *ro, double *pr, double *vx, double *vy, double *by,
              const double *const *g, int ix, int margin, double gm,
              double qav, double safety, double dtmin, double floor_,
              double tend, int nstop, double *t, int *ns, int *merr,
              int *mseed) {
    const double *bx = g[0], *bxm = g[1], *gx = g[2], *gxm = g[3], *sc = g[4], *dsc = g[5];
    const double *scm = g[6], *dscm = g[7], *rr = g[8], *rrm = g[9], *drr = g[10], *drrm = g[11];
    const double *dx = g[12], *dxm = g[13], *x = g[14];
    int nf;
    *merr = 0;
    for (;;) {
        double dt;
        *ns = *ns + 1;
        or1d_cfl_m(&dt, safety, dtmin, merr, gm, bx, ro, pr, vx, vy, by, dx, ix, NULL);
        if (*merr != 0) break;
        *t = *t + dt;
        or1d_mlw_m_bg(ro, pr, vx, vy, by, bx, bxm, dt, qav, gm, gx, gxm, sc, dsc, scm, dscm, rr, rrm, drr, drrm,
                      dx, dxm, ix);
        orsp_trqalfven(vy, dt, *t, x, ix, rr, mseed);
        orsp_bnd(margin, ro, pr, vx, vy, by, ix);
        or1d_chkdav(&nf, ro, vx, floor_, ix);
        or1d_chkdav(&nf, pr, vx, floor_, ix);
        if (!(*ns < nstop && *t < tend)) break;
    }
}
