/*
 * cans_oracle_hlld.c -- C restatement of the HLLD MUSCL-Hancock engine of
 * cansxx-nosupport (Miyoshi & Kusano 2005 flux, mu units) -- test
 * infrastructure only; see cans_oracle.h.  Citations are to
 * cans_beta_vary/cansxx-nosupport/.
 *
 * Faces and temporaries the reference leaves unset are NaN here; they only
 * reach ghost cells, which the caller's boundary fill overwrites.
 */
#include "cans_oracle.h"
#include "cans_oracle_util.h"

/* hdhlld/hlld_flux.f90:51-280, one face; out = fro, fee, frx, fry, frz, fby, fbz */
static void hlld1(double gm, double zfloor, double bxs, double rol, double prl, double vxl, double vyl, double vzl,
                  double byl, double bzl, double ror, double prr, double vxr, double vyr, double vzr, double byr,
                  double bzr, double *o) {
    double bxsq = sq(bxs);
    double pbl = 0.5 * (sq(bxs) + sq(byl) + sq(bzl));
    double ptl = prl + pbl;
    double eel = prl / (gm - 1.0) + 0.5 * rol * (sq(vxl) + sq(vyl) + sq(vzl)) + pbl;
    double rxl = rol * vxl, ryl = rol * vyl, rzl = rol * vzl;
    double pbr = 0.5 * (sq(bxs) + sq(byr) + sq(bzr));
    double ptr = prr + pbr;
    double eer = prr / (gm - 1.0) + 0.5 * (ror * (sq(vxr) + sq(vyr) + sq(vzr))) + pbr;
    double rxr = ror * vxr, ryr = ror * vyr, rzr = ror * vzr;
    double gmpl = gm * prl, gmpr = gm * prr;
    double gpbl = gmpl + 2.0 * pbl, gpbr = gmpr + 2.0 * pbr;
    double cfl = sqrt((gpbl + sqrt(sq(gpbl) - 4.0 * gmpl * bxsq)) / (2.0 * rol));
    double cfr = sqrt((gpbr + sqrt(sq(gpbr) - 4.0 * gmpr * bxsq)) / (2.0 * ror));
    double sl = fminF(vxl, vxr) - fmaxF(cfl, cfr);
    double sr = fmaxF(vxl, vxr) + fmaxF(cfl, cfr);
    double frol = rxl, frxl = rxl * vxl + ptl - bxsq, fryl = rol * vxl * vyl - bxs * byl;
    double frzl = rol * vxl * vzl - bxs * bzl;
    double feel = vxl * (eel + ptl - bxsq) - bxs * (vyl * byl + vzl * bzl);
    double fbyl = byl * vxl - bxs * vyl, fbzl = bzl * vxl - bxs * vzl;
    double fror = rxr, frxr = rxr * vxr + ptr - bxsq, fryr = ror * vxr * vyr - bxs * byr;
    double frzr = ror * vxr * vzr - bxs * bzr;
    double feer = vxr * (eer + ptr - bxsq) - bxs * (vyr * byr + vzr * bzr);
    double fbyr = byr * vxr - bxs * vyr, fbzr = bzr * vxr - bxs * vzr;
    if (sl >= 0.0) {
        o[0] = frol; o[2] = frxl; o[3] = fryl; o[4] = frzl; o[1] = feel; o[5] = fbyl; o[6] = fbzl;
        return;
    }
    if (sr <= 0.0) {
        o[0] = fror; o[2] = frxr; o[3] = fryr; o[4] = frzr; o[1] = feer; o[5] = fbyr; o[6] = fbzr;
        return;
    }
    double sdl = sl - vxl, sdr = sr - vxr;
    double sm = (sdr * ror * vxr - sdl * rol * vxl - ptr + ptl) / (sdr * ror - sdl * rol);
    double sdml = sl - sm, sdmr = sr - sm;
    double ptst = (sdr * ror * ptl - sdl * rol * ptr + rol * ror * sdr * sdl * (vxr - vxl)) / (sdr * ror - sdl * rol);
    /* left star state */
    double temp_fst = (rol * sdl * sdml - bxsq);
    double sign1 = signF(1.0, temp_fst + zfloor), sign2 = signF(1.0, temp_fst - zfloor), sign3 = sign1 * sign2;
    double tmp = bxs * (sdl - sdml) / (temp_fst + fminF(0.0, sign3));
    double rolst = fmaxF(0.0, sign3) * (rol * sdl / sdml) - fminF(0.0, sign3) * rol;
    double vxlst = fmaxF(0.0, sign3) * (sm) - fminF(0.0, sign3) * vxl;
    double rxlst = rolst * vxlst;
    double vylst = fmaxF(0.0, sign3) * (vyl - byl * tmp) - fminF(0.0, sign3) * vyl;
    double rylst = rolst * vylst;
    double vzlst = fmaxF(0.0, sign3) * (vzl - bzl * tmp) - fminF(0.0, sign3) * vzl;
    double rzlst = rolst * vzlst;
    tmp = (rol * sq(sdl) - bxsq) / (temp_fst + fminF(0.0, sign3));
    double bylst = fmaxF(0.0, sign3) * (byl * tmp) - fminF(0.0, sign3) * byl;
    double bzlst = fmaxF(0.0, sign3) * (bzl * tmp) - fminF(0.0, sign3) * bzl;
    double vdbstl = (vxlst * bxs + vylst * bylst + vzlst * bzlst);
    double eelst = fmaxF(0.0, sign3) * ((sdl * eel - ptl * vxl + ptst * sm + bxs * (vxl * bxs + vyl * byl + vzl * bzl - vdbstl)) / sdml)
                 - fminF(0.0, sign3) * eel;
    /* right star state */
    temp_fst = (ror * sdr * sdmr - bxsq);
    sign1 = signF(1.0, temp_fst + zfloor); sign2 = signF(1.0, temp_fst - zfloor); sign3 = sign1 * sign2;
    tmp = bxs * (sdr - sdmr) / (temp_fst + fminF(0.0, sign3));
    double rorst = fmaxF(0.0, sign3) * (ror * sdr / sdmr) - fminF(0.0, sign3) * ror;
    double vxrst = fmaxF(0.0, sign3) * (sm) - fminF(0.0, sign3) * vxr;
    double rxrst = rorst * vxrst;
    double vyrst = fmaxF(0.0, sign3) * (vyr - byr * tmp) - fminF(0.0, sign3) * vyr;
    double ryrst = rorst * vyrst;
    double vzrst = fmaxF(0.0, sign3) * (vzr - bzr * tmp) - fminF(0.0, sign3) * vzr;
    double rzrst = rorst * vzrst;
    tmp = (ror * sq(sdr) - bxsq) / (temp_fst + fminF(0.0, sign3));
    double byrst = fmaxF(0.0, sign3) * (byr * tmp) - fminF(0.0, sign3) * byr;
    double bzrst = fmaxF(0.0, sign3) * (bzr * tmp) - fminF(0.0, sign3) * bzr;
    double vdbstr = (vxrst * bxs + vyrst * byrst + vzrst * bzrst);
    double eerst = fmaxF(0.0, sign3) * ((sdr * eer - ptr * vxr + ptst * sm + bxs * (vxr * bxs + vyr * byr + vzr * bzr - vdbstr)) / sdmr)
                 - fminF(0.0, sign3) * eer;
    double sqrtrol = sqrt(rolst), sqrtror = sqrt(rorst);
    double slst = sm - fabs(bxs) / sqrtrol;
    double srst = sm + fabs(bxs) / sqrtror;
    double roldst, rxldst, ryldst, rzldst, eeldst, byldst, bzldst;
    double rordst, rxrdst, ryrdst, rzrdst, eerdst, byrdst, bzrdst;
    if (bxs == 0.0) {
        roldst = rolst; rxldst = rxlst; ryldst = rylst; rzldst = rzlst; eeldst = eelst; byldst = bylst; bzldst = bzlst;
        rordst = rorst; rxrdst = rxrst; ryrdst = ryrst; rzrdst = rzrst; eerdst = eerst; byrdst = byrst; bzrdst = bzrst;
    } else {
        double invsumro = 1.0 / (sqrtrol + sqrtror);
        double signbx = signF(1.0, bxs);
        roldst = rolst; rordst = rorst; rxldst = rxlst; rxrdst = rxrst;
        double temp = invsumro * (sqrtrol * vylst + sqrtror * vyrst + signbx * (byrst - bylst));
        double vyldst = temp;
        ryldst = roldst * temp; ryrdst = rordst * temp;
        temp = invsumro * (sqrtrol * vzlst + sqrtror * vzrst + signbx * (bzrst - bzlst));
        double vzldst = temp;
        rzldst = roldst * temp; rzrdst = rordst * temp;
        temp = invsumro * (sqrtrol * byrst + sqrtror * bylst + signbx * sqrtrol * sqrtror * (vyrst - vylst));
        byldst = temp; byrdst = temp;
        temp = invsumro * (sqrtrol * bzrst + sqrtror * bzlst + signbx * sqrtrol * sqrtror * (vzrst - vzlst));
        bzldst = temp; bzrdst = temp;
        temp = sm * bxs + vyldst * byldst + vzldst * bzldst;
        eeldst = eelst - sqrtrol * signbx * (vdbstl - temp);
        eerdst = eerst + sqrtror * signbx * (vdbstr - temp);
    }
    if (slst >= 0.0) {
        o[0] = frol + sl * (rolst - rol); o[2] = frxl + sl * (rxlst - rxl); o[3] = fryl + sl * (rylst - ryl);
        o[4] = frzl + sl * (rzlst - rzl); o[1] = feel + sl * (eelst - eel); o[5] = fbyl + sl * (bylst - byl);
        o[6] = fbzl + sl * (bzlst - bzl);
    } else if (sm >= 0.0) {
        double temp = slst - sl;
        o[0] = frol - sl * rol - temp * rolst + slst * roldst;
        o[2] = frxl - sl * rxl - temp * rxlst + slst * rxldst;
        o[3] = fryl - sl * ryl - temp * rylst + slst * ryldst;
        o[4] = frzl - sl * rzl - temp * rzlst + slst * rzldst;
        o[1] = feel - sl * eel - temp * eelst + slst * eeldst;
        o[5] = fbyl - sl * byl - temp * bylst + slst * byldst;
        o[6] = fbzl - sl * bzl - temp * bzlst + slst * bzldst;
    } else if (srst > 0.0) {
        double temp = srst - sr;
        o[0] = fror - sr * ror - temp * rorst + srst * rordst;
        o[2] = frxr - sr * rxr - temp * rxrst + srst * rxrdst;
        o[3] = fryr - sr * ryr - temp * ryrst + srst * ryrdst;
        o[4] = frzr - sr * rzr - temp * rzrst + srst * rzrdst;
        o[1] = feer - sr * eer - temp * eerst + srst * eerdst;
        o[5] = fbyr - sr * byr - temp * byrst + srst * byrdst;
        o[6] = fbzr - sr * bzr - temp * bzrst + srst * bzrdst;
    } else {
        o[0] = fror + sr * (rorst - ror); o[2] = frxr + sr * (rxrst - rxr); o[3] = fryr + sr * (ryrst - ryr);
        o[4] = frzr + sr * (rzrst - rzr); o[1] = feer + sr * (eerst - eer); o[5] = fbyr + sr * (byrst - byr);
        o[6] = fbzr + sr * (bzrst - bzr);
    }
}

enum { P_RO, P_PR, P_VX, P_VY, P_VZ, P_BX, P_BY, P_BZ, P_N };
enum { U_RO, U_EE, U_RX, U_RY, U_RZ, U_BX, U_BY, U_BZ, U_N };

static inline double flmtH(double a, double b) {
    return fmaxF(0.0, fminF(b * signF(1.0, a), fabs(a))) * signF(1.0, a);
}

/* hdhlld/hlld_mh.f:2-127 for one direction; prim = rof..bzf, un = ron..bzn,
 * w0/w1 and bb/f are the automatic arrays of the reference (kept across
 * directions as the reference's are) */
static void hlld_mh_dir(double *const *prim, double *const *un, double gm, double dt, const double *dx,
                        const double *dy, const double *dz, int ix, int jx, int kx, int mdir, int mstage,
                        double *const *w0, double *const *w1, double *bb, double *const *f) {
    const size_t s = mdir == 1 ? 1 : (mdir == 2 ? (size_t)ix : (size_t)ix * jx);
    const int ni = ix - (mdir == 1), nj = jx - (mdir == 2), nk = kx - (mdir == 3);
    const int nd = mdir == 1 ? ix : (mdir == 2 ? jx : kx);
    /* interpolation (interpolation/inp_const.f, inp_tvdminmod.f) */
    for (int q = 0; q < P_N; q++)
        for (int k = 1; k <= nk; k++)
            for (int j = 1; j <= nj; j++)
                for (int i = 1; i <= ni; i++) {
                    size_t p = I3(i, j, k);
                    int a = mdir == 1 ? i : (mdir == 2 ? j : k);
                    const double *da = prim[q];
                    if (mstage == 1 || a == 1 || a == nd - 1) {
                        w0[q][p] = da[p];
                        w1[q][p] = da[p + s];
                    } else {
                        w0[q][p] = da[p] + 0.5 * flmtH(da[p] - da[p - s], da[p + s] - da[p]);
                        w1[q][p] = da[p + s] - 0.5 * flmtH(da[p + 2 * s] - da[p + s], da[p + s] - da[p]);
                    }
                }
    /* face-normal field and the flux (hlld_mh.f:43-95) */
    const double *bn = prim[mdir == 1 ? P_BX : (mdir == 2 ? P_BY : P_BZ)];
    for (int k = 1; k <= nk; k++)
        for (int j = 1; j <= nj; j++)
            for (int i = 1; i <= ni; i++) {
                size_t p = I3(i, j, k);
                bb[p] = (bn[p] + bn[p + s]) / 2.0;
            }
    int vn, v1, v2, b1, b2, fn, f1, f2, g1, g2, gn;
    if (mdir == 1) { vn = P_VX; v1 = P_VY; v2 = P_VZ; b1 = P_BY; b2 = P_BZ; fn = U_RX; f1 = U_RY; f2 = U_RZ; g1 = U_BY; g2 = U_BZ; gn = U_BX; }
    else if (mdir == 2) { vn = P_VY; v1 = P_VZ; v2 = P_VX; b1 = P_BZ; b2 = P_BX; fn = U_RY; f1 = U_RZ; f2 = U_RX; g1 = U_BZ; g2 = U_BX; gn = U_BY; }
    else { vn = P_VZ; v1 = P_VX; v2 = P_VY; b1 = P_BX; b2 = P_BY; fn = U_RZ; f1 = U_RX; f2 = U_RY; g1 = U_BX; g2 = U_BY; gn = U_BZ; }
    const double floor_ = 1.0e-10;
    for (int k = 1; k <= kx; k++)      /* hlld_flux.f90 loops over every i,j,k */
        for (int j = 1; j <= jx; j++)
            for (int i = 1; i <= ix; i++) {
                size_t p = I3(i, j, k);
                double o[7];
                hlld1(gm, floor_, bb[p], w0[P_RO][p], w0[P_PR][p], w0[vn][p], w0[v1][p], w0[v2][p], w0[b1][p], w0[b2][p],
                      w1[P_RO][p], w1[P_PR][p], w1[vn][p], w1[v1][p], w1[v2][p], w1[b1][p], w1[b2][p], o);
                f[U_RO][p] = o[0]; f[U_EE][p] = o[1]; f[fn][p] = o[2]; f[f1][p] = o[3]; f[f2][p] = o[4];
                f[g1][p] = o[5]; f[g2][p] = o[6];
            }
    /* the normal-field flux is zeroed on i=1..ix-1 (hlld_mh.f:56-94) */
    for (int k = 1; k <= kx; k++)
        for (int j = 1; j <= jx; j++)
            for (int i = 1; i <= ix - 1; i++) f[gn][I3(i, j, k)] = 0.0;
    /* common/adddfdx.f */
    const double *dd = mdir == 1 ? dx : (mdir == 2 ? dy : dz);
    for (int q = 0; q < U_N; q++)
        for (int k = (mdir == 3 ? 2 : 1); k <= (mdir == 3 ? kx - 1 : kx); k++)
            for (int j = (mdir == 2 ? 2 : 1); j <= (mdir == 2 ? jx - 1 : jx); j++)
                for (int i = (mdir == 1 ? 2 : 1); i <= (mdir == 1 ? ix - 1 : ix); i++) {
                    size_t p = I3(i, j, k);
                    int a = mdir == 1 ? i : (mdir == 2 ? j : k);
                    un[q][p] = un[q][p] - dt * ((f[q][p] - f[q][p - s]) / dd[I1(a)]);
                }
}

/* md_mhdshktb/hdhlld-x/engine.f:2-73 */
void orh_engine(double *ro, double *pr, double *vx, double *vy, double *vz,
                double *bx, double *by, double *bz, double gm, double mu,
                double dt, const double *dx, const double *dy,
                const double *dz, int ix, int jx, int kx, const int *mfdim) {
    size_t n = (size_t)ix * jx * kx;
    double *prim[P_N] = {ro, pr, vx, vy, vz, bx, by, bz};
    double *u[U_N], *uh[U_N], *u2[U_N], *ph[P_N], *w0[P_N], *w1[P_N], *f[U_N];
    for (int q = 0; q < U_N; q++) { u[q] = dalloc(n); uh[q] = dalloc(n); u2[q] = dalloc(n); f[q] = dalloc(n); }
    for (int q = 0; q < P_N; q++) { ph[q] = dalloc(n); w0[q] = dalloc(n); w1[q] = dalloc(n); }
    double *bb = dalloc(n), *eh = dalloc(n);
    for (size_t p = 0; p < n; p++) {
        u[U_RO][p] = ro[p];
        u[U_RX][p] = vx[p] * ro[p];       /* common/vvtorv.f */
        u[U_RY][p] = vy[p] * ro[p];
        u[U_RZ][p] = vz[p] * ro[p];
        eh[p] = pr[p] / (gm - 1.0);       /* common/pr2eh.f */
        double ek = (sq(vx[p]) + sq(vy[p]) + sq(vz[p])) * ro[p] / 2.0;   /* common/eh2ee.f */
        double em = (sq(bx[p]) + sq(by[p]) + sq(bz[p])) / (2.0 * mu);
        u[U_EE][p] = eh[p] + ek + em;
        u[U_BX][p] = bx[p]; u[U_BY][p] = by[p]; u[U_BZ][p] = bz[p];
    }
    for (int q = 0; q < U_N; q++)         /* hdroe/roereset.f */
        for (size_t p = 0; p < n; p++) { uh[q][p] = u[q][p]; u2[q][p] = u[q][p]; }
    double dth = 0.5 * dt;
    for (int d = 1; d <= 3; d++)
        if (mfdim[d - 1] == 1) hlld_mh_dir(prim, uh, gm, dth, dx, dy, dz, ix, jx, kx, d, 1, w0, w1, bb, f);
    for (size_t p = 0; p < n; p++) {
        ph[P_RO][p] = uh[U_RO][p];
        ph[P_VX][p] = uh[U_RX][p] / uh[U_RO][p];   /* common/rvtovv.f */
        ph[P_VY][p] = uh[U_RY][p] / uh[U_RO][p];
        ph[P_VZ][p] = uh[U_RZ][p] / uh[U_RO][p];
        ph[P_BX][p] = uh[U_BX][p]; ph[P_BY][p] = uh[U_BY][p]; ph[P_BZ][p] = uh[U_BZ][p];
        double ek = (sq(ph[P_VX][p]) + sq(ph[P_VY][p]) + sq(ph[P_VZ][p])) * ph[P_RO][p] / 2.0;   /* common/ee2eh.f */
        double em = (sq(ph[P_BX][p]) + sq(ph[P_BY][p]) + sq(ph[P_BZ][p])) / (2.0 * mu);
        double ehh = uh[U_EE][p] - ek - em;
        ph[P_PR][p] = (gm - 1.0) * ehh;            /* common/eh2pr.f */
    }
    for (int d = 1; d <= 3; d++)
        if (mfdim[d - 1] == 1) hlld_mh_dir(ph, u2, gm, dt, dx, dy, dz, ix, jx, kx, d, 2, w0, w1, bb, f);
    for (size_t p = 0; p < n; p++) {           /* common/daupdate.f, then primitives */
        ro[p] = u2[U_RO][p];
        bx[p] = u2[U_BX][p]; by[p] = u2[U_BY][p]; bz[p] = u2[U_BZ][p];
        vx[p] = u2[U_RX][p] / ro[p];
        vy[p] = u2[U_RY][p] / ro[p];
        vz[p] = u2[U_RZ][p] / ro[p];
        double ek = (sq(vx[p]) + sq(vy[p]) + sq(vz[p])) * ro[p] / 2.0;
        double em = (sq(bx[p]) + sq(by[p]) + sq(bz[p])) / (2.0 * mu);
        double ehf = u2[U_EE][p] - ek - em;
        pr[p] = (gm - 1.0) * ehf;
    }
    for (int q = 0; q < U_N; q++) { free(u[q]); free(uh[q]); free(u2[q]); free(f[q]); }
    for (int q = 0; q < P_N; q++) { free(ph[q]); free(w0[q]); free(w1[q]); }
    free(bb); free(eh);
}

/* ---- the helper routines themselves, with the reference argument lists */
/* hdhlld/hlld_flux.f90:2-3, every face i,j,k (:141-143); w arrays (ix,jx,kx,2) */
void orh_hlld_flux(const double *row, const double *prw, const double *vxw, const double *vyw, const double *vzw,
                   const double *bx, const double *byw, const double *bzw, double gm, int ix, int jx, int kx,
                   double zfloor, double *fro, double *fee, double *frx, double *fry, double *frz, double *fby,
                   double *fbz) {
    const size_t n = (size_t)ix * jx * kx;
    for (size_t p = 0; p < n; p++) {
        const size_t q = p + n;
        double o[7];
        hlld1(gm, zfloor, bx[p], row[p], prw[p], vxw[p], vyw[p], vzw[p], byw[p], bzw[p], row[q], prw[q], vxw[q],
              vyw[q], vzw[q], byw[q], bzw[q], o);
        fro[p] = o[0]; fee[p] = o[1]; frx[p] = o[2]; fry[p] = o[3]; frz[p] = o[4]; fby[p] = o[5]; fbz[p] = o[6];
    }
}
/* hdhlld/hlld_mh.f:2-4: the directions with mfdim = 1, in order, on one set
 * of automatic arrays (NaN where the reference leaves them unset) */
void orh_hlld_mh(double *rof, double *prf, double *vxf, double *vyf, double *vzf, double *bxf, double *byf,
                 double *bzf, double *ron, double *een, double *rxn, double *ryn, double *rzn, double *bxn,
                 double *byn, double *bzn, double gm, double dt, const double *dx, const double *dy,
                 const double *dz, int ix, int jx, int kx, const int *mfdim, int mstage) {
    size_t n = (size_t)ix * jx * kx;
    double *prim[P_N] = {rof, prf, vxf, vyf, vzf, bxf, byf, bzf};
    double *un[U_N] = {ron, een, rxn, ryn, rzn, bxn, byn, bzn};
    double *w0[P_N], *w1[P_N], *f[U_N];
    for (int q = 0; q < P_N; q++) { w0[q] = dalloc(n); w1[q] = dalloc(n); }
    for (int q = 0; q < U_N; q++) f[q] = dalloc(n);
    double *bb = dalloc(n);
    for (int d = 1; d <= 3; d++)
        if (mfdim[d - 1] == 1) hlld_mh_dir(prim, un, gm, dt, dx, dy, dz, ix, jx, kx, d, mstage, w0, w1, bb, f);
    for (int q = 0; q < P_N; q++) { free(w0[q]); free(w1[q]); }
    for (int q = 0; q < U_N; q++) free(f[q]);
    free(bb);
}
