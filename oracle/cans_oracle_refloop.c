/*
 * cans_oracle_refloop.c -- the CANS main-program loop driving routines passed
 * in as function pointers (test infrastructure only).
 *
 * The CPU baselines time the reference routines built from their own
 * sources (oracle/_ref); calling them once per step from Python would add
 * the interpreter's call overhead to a 1024-cell step that takes a few
 * microseconds.  This loop calls them directly with the Fortran convention
 * (every argument by reference), in the order of
 * cans1d/md_mhdshktb/hdmlw/main.f:132-186:
 *     ns=ns+1; cfl_m; stop on merr; t=t+dt; mlw_m; bnd; loop while
 *     ns < nstop and t < tend
 * with bnd = bdfrex at both ends on ro, pr, vx, vy, by (the free-boundary
 * shock tube).
 */
typedef void (*cfl1_fn)(double *dt, const double *safety, const double *dtmin, int *merr, const double *gm,
                        const double *bx, const double *ro, const double *pr, const double *vx, const double *vy,
                        const double *by, const double *dx, const int *ix);
typedef void (*mlw1_fn)(double *ro, double *pr, double *vx, double *vy, double *by, const double *bx,
                        const double *bxm, const double *dt, const double *qav, const double *gm, const double *dx,
                        const double *dxm, const int *ix);
typedef void (*bd1_fn)(const int *mbnd, const int *margin, double *qq, const int *ix);

/* Returns the loop entries ns; *t_out the final time, *merr_out cfl_m's merr. */
int orref_loop_1d(void *cfl, void *mlw, void *bdfre, double *ro, double *pr, double *vx, double *vy, double *by,
                  const double *bx, const double *bxm, const double *dx, const double *dxm, int ix, int margin,
                  double gm, double safety, double qav, double dtmin, double tend, int nstop, double *t_out,
                  int *merr_out) {
    cfl1_fn c = (cfl1_fn)cfl;
    mlw1_fn m = (mlw1_fn)mlw;
    bd1_fn b = (bd1_fn)bdfre;
    double t = 0.0, dt = 0.0;
    int ns = 0, merr = 0;
    const int lo = 0, hi = 1;
    double *f[5] = {ro, pr, vx, vy, by};
    do {
        ns = ns + 1;
        c(&dt, &safety, &dtmin, &merr, &gm, bx, ro, pr, vx, vy, by, dx, &ix);
        if (merr != 0) break;
        t = t + dt;
        m(ro, pr, vx, vy, by, bx, bxm, &dt, &qav, &gm, dx, dxm, &ix);
        for (int q = 0; q < 5; q++) {
            b(&lo, &margin, f[q], &ix);
            b(&hi, &margin, f[q], &ix);
        }
    } while (ns < nstop && t < tend);
    *t_out = t;
    *merr_out = merr;
    return ns;
}
