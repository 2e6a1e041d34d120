// cans_hlld.hip -- HLLD MUSCL-Hancock MHD engine (cansxx-nosupport/hdhlld,
// Miyoshi & Kusano 2005; mu units) on gfx950, FP64 (bit-exact) and FP32-flux
// mixed precision.
//
// engine.f:2-73 as kernels: per stage and direction one sweep kernel that
// computes the face fluxes (inp_const / inp_tvdminmod + hlld_flux.f90) and
// applies them to the cells (adddfdx.f).  The face fluxes never go to
// memory: along x a wave covers 64 faces and takes the west face of each
// cell from the neighbouring lane (DPP); along y and z each thread marches
// over a segment of cells with the four stencil cells of its face in a
// sliding register window and carries the previous face (one face per
// segment is computed twice, by the same operations); the next cell's
// window values and conserved state are loaded one face ahead (with FP64
// fluxes the conserved state goes through LDS by global->LDS loads, as
// the registers are full).  The first (x) sweep
// of a stage also sets up the conserved state from the primitives and the
// last (z) sweep recovers the primitives (the separate hlld_cons/hlld_prim
// passes remain for runs without an x or z sweep).  The state accumulates
// in FP64.  In a periodic run the stage-1 x sweep also advances the device
// clock and the stage-2 z sweep takes the next step's cfl_m (image_min,
// cfl3_term), so a step needs no cfl pass or clock kernel of its own.
#include "cans_common.hpp"
#include "cans_clock.hpp"
#include "cans_roe.hpp"
#include "cans_roeflux.hpp"
#include "cans_wave.hpp"
#include "../../include/cans_gpu.h"

namespace cans {
namespace {

enum { P_RO, P_PR, P_VX, P_VY, P_VZ, P_BX, P_BY, P_BZ };
enum { U_RO, U_EE, U_RX, U_RY, U_RZ, U_BX, U_BY, U_BZ };

// Run flag / dt of the step: the clock as the step's clock kernel (or, with
// `enter`, the step's first kernel here, block 0 writing it) advanced it.
CANS_DEV bool hstep(const HlldArgs &a, double &dt, bool enter = false) {
    if (!a.clk) { dt = a.dt; return true; }
    const StepCtl c = clock_use(a.clk, a.step, !enter, blockIdx.x == 0 && threadIdx.x == 0);
    dt = c.dt;
    return c.run;
}

// hdhlld/hlld_flux.f90:51-280 for one face; o = fro, fee, frx, fry, frz, fby, fbz
template <class T>
CANS_DEV void hlld_face(T gm, T zfloor, T bxs, T rol, T prl, T vxl, T vyl, T vzl, T byl, T bzl, T ror, T prr,
                        T vxr, T vyr, T vzr, T byr, T bzr, T *o) {
    const T one = T(1), half = T(0.5), zero = T(0), two = T(2);
    const T bxsq = bxs * bxs;
    const T pbl = half * (bxs * bxs + byl * byl + bzl * bzl);
    const T ptl = prl + pbl;
    const T eel = prl / (gm - one) + half * rol * (vxl * vxl + vyl * vyl + vzl * vzl) + pbl;
    const T rxl = rol * vxl, ryl = rol * vyl, rzl = rol * vzl;
    const T pbr = half * (bxs * bxs + byr * byr + bzr * bzr);
    const T ptr = prr + pbr;
    const T eer = prr / (gm - one) + half * (ror * (vxr * vxr + vyr * vyr + vzr * vzr)) + pbr;
    const T rxr = ror * vxr, ryr = ror * vyr, rzr = ror * vzr;
    const T gmpl = gm * prl, gmpr = gm * prr;
    const T gpbl = gmpl + two * pbl, gpbr = gmpr + two * pbr;
    const T cfl = tsqrt((gpbl + tsqrt(gpbl * gpbl - T(4) * gmpl * bxsq)) / (two * rol));
    const T cfr = tsqrt((gpbr + tsqrt(gpbr * gpbr - T(4) * gmpr * bxsq)) / (two * ror));
    const T sl = fminF(vxl, vxr) - fmaxF(cfl, cfr);
    const T sr = fmaxF(vxl, vxr) + fmaxF(cfl, cfr);
    const T frol = rxl, frxl = rxl * vxl + ptl - bxsq, fryl = rol * vxl * vyl - bxs * byl;
    const T frzl = rol * vxl * vzl - bxs * bzl;
    const T feel = vxl * (eel + ptl - bxsq) - bxs * (vyl * byl + vzl * bzl);
    const T fbyl = byl * vxl - bxs * vyl, fbzl = bzl * vxl - bxs * vzl;
    const T fror = rxr, frxr = rxr * vxr + ptr - bxsq, fryr = ror * vxr * vyr - bxs * byr;
    const T frzr = ror * vxr * vzr - bxs * bzr;
    const T feer = vxr * (eer + ptr - bxsq) - bxs * (vyr * byr + vzr * bzr);
    const T fbyr = byr * vxr - bxs * vyr, fbzr = bzr * vxr - bxs * vzr;
    if (sl >= zero) {
        o[0] = frol; o[1] = feel; o[2] = frxl; o[3] = fryl; o[4] = frzl; o[5] = fbyl; o[6] = fbzl;
        return;
    }
    if (sr <= zero) {
        o[0] = fror; o[1] = feer; o[2] = frxr; o[3] = fryr; o[4] = frzr; o[5] = fbyr; o[6] = fbzr;
        return;
    }
    const T sdl = sl - vxl, sdr = sr - vxr;
    const T den = sdr * ror - sdl * rol;
    const T sm = (sdr * ror * vxr - sdl * rol * vxl - ptr + ptl) / den;
    const T sdml = sl - sm, sdmr = sr - sm;
    const T ptst = (sdr * ror * ptl - sdl * rol * ptr + rol * ror * sdr * sdl * (vxr - vxl)) / den;
    // left star state
    T tf = (rol * sdl * sdml - bxsq);
    T s3 = signF(one, tf + zfloor) * signF(one, tf - zfloor);
    T mx = fmaxF(zero, s3), mn = fminF(zero, s3);
    T tmp = bxs * (sdl - sdml) / (tf + mn);
    const T rolst = mx * (rol * sdl / sdml) - mn * rol;
    const T vxlst = mx * (sm) - mn * vxl;
    const T rxlst = rolst * vxlst;
    const T vylst = mx * (vyl - byl * tmp) - mn * vyl;
    const T rylst = rolst * vylst;
    const T vzlst = mx * (vzl - bzl * tmp) - mn * vzl;
    const T rzlst = rolst * vzlst;
    tmp = (rol * (sdl * sdl) - bxsq) / (tf + mn);
    const T bylst = mx * (byl * tmp) - mn * byl;
    const T bzlst = mx * (bzl * tmp) - mn * bzl;
    const T vdbstl = (vxlst * bxs + vylst * bylst + vzlst * bzlst);
    const T eelst = mx * ((sdl * eel - ptl * vxl + ptst * sm + bxs * (vxl * bxs + vyl * byl + vzl * bzl - vdbstl)) / sdml)
                  - mn * eel;
    // right star state
    tf = (ror * sdr * sdmr - bxsq);
    s3 = signF(one, tf + zfloor) * signF(one, tf - zfloor);
    mx = fmaxF(zero, s3);
    mn = fminF(zero, s3);
    tmp = bxs * (sdr - sdmr) / (tf + mn);
    const T rorst = mx * (ror * sdr / sdmr) - mn * ror;
    const T vxrst = mx * (sm) - mn * vxr;
    const T rxrst = rorst * vxrst;
    const T vyrst = mx * (vyr - byr * tmp) - mn * vyr;
    const T ryrst = rorst * vyrst;
    const T vzrst = mx * (vzr - bzr * tmp) - mn * vzr;
    const T rzrst = rorst * vzrst;
    tmp = (ror * (sdr * sdr) - bxsq) / (tf + mn);
    const T byrst = mx * (byr * tmp) - mn * byr;
    const T bzrst = mx * (bzr * tmp) - mn * bzr;
    const T vdbstr = (vxrst * bxs + vyrst * byrst + vzrst * bzrst);
    const T eerst = mx * ((sdr * eer - ptr * vxr + ptst * sm + bxs * (vxr * bxs + vyr * byr + vzr * bzr - vdbstr)) / sdmr)
                  - mn * eer;
    const T sqrtrol = tsqrt(rolst), sqrtror = tsqrt(rorst);
    const T slst = sm - tabs(bxs) / sqrtrol;
    const T srst = sm + tabs(bxs) / sqrtror;
    T roldst, rxldst, ryldst, rzldst, eeldst, byldst, bzldst;
    T rordst, rxrdst, ryrdst, rzrdst, eerdst, byrdst, bzrdst;
    if (bxs == zero) {
        roldst = rolst; rxldst = rxlst; ryldst = rylst; rzldst = rzlst; eeldst = eelst; byldst = bylst; bzldst = bzlst;
        rordst = rorst; rxrdst = rxrst; ryrdst = ryrst; rzrdst = rzrst; eerdst = eerst; byrdst = byrst; bzrdst = bzrst;
    } else {
        const T invsumro = one / (sqrtrol + sqrtror);
        const T signbx = signF(one, bxs);
        roldst = rolst; rordst = rorst; rxldst = rxlst; rxrdst = rxrst;
        T temp = invsumro * (sqrtrol * vylst + sqrtror * vyrst + signbx * (byrst - bylst));
        const T vyldst = temp;
        ryldst = roldst * temp; ryrdst = rordst * temp;
        temp = invsumro * (sqrtrol * vzlst + sqrtror * vzrst + signbx * (bzrst - bzlst));
        const T vzldst = temp;
        rzldst = roldst * temp; rzrdst = rordst * temp;
        temp = invsumro * (sqrtrol * byrst + sqrtror * bylst + signbx * sqrtrol * sqrtror * (vyrst - vylst));
        byldst = temp; byrdst = temp;
        temp = invsumro * (sqrtrol * bzrst + sqrtror * bzlst + signbx * sqrtrol * sqrtror * (vzrst - vzlst));
        bzldst = temp; bzrdst = temp;
        temp = sm * bxs + vyldst * byldst + vzldst * bzldst;
        eeldst = eelst - sqrtrol * signbx * (vdbstl - temp);
        eerdst = eerst + sqrtror * signbx * (vdbstr - temp);
    }
    if (slst >= zero) {
        o[0] = frol + sl * (rolst - rol); o[2] = frxl + sl * (rxlst - rxl); o[3] = fryl + sl * (rylst - ryl);
        o[4] = frzl + sl * (rzlst - rzl); o[1] = feel + sl * (eelst - eel); o[5] = fbyl + sl * (bylst - byl);
        o[6] = fbzl + sl * (bzlst - bzl);
    } else if (sm >= zero) {
        const T temp = slst - sl;
        o[0] = frol - sl * rol - temp * rolst + slst * roldst;
        o[2] = frxl - sl * rxl - temp * rxlst + slst * rxldst;
        o[3] = fryl - sl * ryl - temp * rylst + slst * ryldst;
        o[4] = frzl - sl * rzl - temp * rzlst + slst * rzldst;
        o[1] = feel - sl * eel - temp * eelst + slst * eeldst;
        o[5] = fbyl - sl * byl - temp * bylst + slst * byldst;
        o[6] = fbzl - sl * bzl - temp * bzlst + slst * bzldst;
    } else if (srst > zero) {
        const T temp = srst - sr;
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

CANS_DEV double flmtH(double a, double b) {
    return fmaxF(0.0, fminF(b * signF(1.0, a), __builtin_fabs(a))) * signF(1.0, a);
}

// conserved state of cell p from the primitives (vvtorv.f, pr2eh.f, eh2ee.f)
CANS_DEV void hlld_cons_of(const HlldArgs &a, size_t p, double *U) {
    const double ro = a.prim[P_RO][p], pr = a.prim[P_PR][p], vx = a.prim[P_VX][p], vy = a.prim[P_VY][p];
    const double vz = a.prim[P_VZ][p], bx = a.prim[P_BX][p], by = a.prim[P_BY][p], bz = a.prim[P_BZ][p];
    const double eh = pr / (a.gm - 1.0);
    const double ek = (vx * vx + vy * vy + vz * vz) * ro / 2.0;
    const double em = (bx * bx + by * by + bz * bz) / (2.0 * a.mu);
    U[0] = ro; U[1] = eh + ek + em; U[2] = vx * ro; U[3] = vy * ro; U[4] = vz * ro; U[5] = bx; U[6] = by; U[7] = bz;
}

// conserved state of every cell: vvtorv.f, pr2eh.f, eh2ee.f; into u1 and u2
__global__ __launch_bounds__(256) void hlld_cons(HlldArgs a) {
    double dt;
    if (!hstep(a, dt)) return;
    const size_t N = (size_t)a.ix * a.jx * a.kx;
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    double U[8];
    hlld_cons_of(a, p, U);
    for (int q = 0; q < 8; q++) { a.u1[q][p] = U[q]; a.u2[q][p] = U[q]; }
}

// Flux of the face between cell p and cell p+s along `dir` (face index
// aidx in [1, n_dir-1]), in the U slots ro ee rx ry rz bx by bz (the normal
// field's flux is 0, hlld_mh.f:56-94).
// Face flux from the values of the four cells around it (d0..d3 along the
// sweep, the face between d1 and d2): MUSCL-minmod states where muscl
// (inp_tvdminmod.f), else the piecewise-constant ones (inp_const.f, d0 and
// d3 unused).
template <class T>
CANS_DEV void hlld_face_w(const HlldArgs &a, int dir, bool muscl, const double *d0, const double *d1,
                          const double *d2, const double *d3, double *F) {
    T L[8], R[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        L[q] = muscl ? T(d1[q] + 0.5 * flmtH(d1[q] - d0[q], d2[q] - d1[q])) : T(d1[q]);
        R[q] = muscl ? T(d2[q] - 0.5 * flmtH(d3[q] - d2[q], d2[q] - d1[q])) : T(d2[q]);
    }
    const T bb = T((d1[P_BX + dir] + d2[P_BX + dir]) / 2.0);   // hlld_mh.f:47/65/82
    const int vn = P_VX + dir, v1 = P_VX + (dir + 1) % 3, v2 = P_VX + (dir + 2) % 3;
    const int b1 = P_BX + (dir + 1) % 3, b2 = P_BX + (dir + 2) % 3;
    T o[7];
    hlld_face<T>(T(a.gm), T(1.0e-10), bb, L[P_RO], L[P_PR], L[vn], L[v1], L[v2], L[b1], L[b2], R[P_RO], R[P_PR],
                 R[vn], R[v1], R[v2], R[b1], R[b2], o);
    F[U_RO] = double(o[0]);
    F[U_EE] = double(o[1]);
    F[U_RX + dir] = double(o[2]);
    F[U_RX + (dir + 1) % 3] = double(o[3]);
    F[U_RX + (dir + 2) % 3] = double(o[4]);
    F[U_BX + (dir + 1) % 3] = double(o[5]);
    F[U_BX + (dir + 2) % 3] = double(o[6]);
    F[U_BX + dir] = 0.0;
}

// Face aidx (of ndir along the sweep) between cell p and p+s.
template <class T>
CANS_DEV void hlld_face_flux(const HlldArgs &a, const double *const *P, int dir, int stage, int aidx, int ndir,
                             size_t p, size_t s, double *F) {
    const bool muscl = stage == 2 && aidx >= 2 && aidx <= ndir - 2;   // inp_tvdminmod.f
    // four cells, clamped to the face's two where the states are piecewise
    // constant (inp_const.f)
    const size_t pm = muscl ? p - s : p, p3 = muscl ? p + 2 * s : p + s;
    double d0[8], d1[8], d2[8], d3[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const double *da = P[q];
        d0[q] = da[pm]; d1[q] = da[p]; d2[q] = da[p + s]; d3[q] = da[p3];
    }
    hlld_face_w<T>(a, dir, muscl, d0, d1, d2, d3, F);
}

// primitives of one cell from its conserved state (rvtovv.f, ee2eh.f, eh2pr.f);
// with `cfl`, returns the next step's cfl_m term of the cell (dg: the
// smallest spacing among its periodic images; the sums v2 and b2 are those
// of cfl_m), else 1.e20
CANS_DEV double hlld_prim_of(const HlldArgs &a, const double *U, double *const *P, size_t p, bool cfl = false,
                             double dg = 0.0) {
    const double ro = U[U_RO];
    const double vx = U[U_RX] / ro;
    const double vy = U[U_RY] / ro;
    const double vz = U[U_RZ] / ro;
    const double bx = U[U_BX], by = U[U_BY], bz = U[U_BZ];
    const double v2 = vx * vx + vy * vy + vz * vz;
    const double b2 = bx * bx + by * by + bz * bz;
    const double ek = v2 * ro / 2.0;
    const double em = b2 / (2.0 * a.mu);
    const double eh = U[U_EE] - ek - em;
    const double pr = (a.gm - 1.0) * eh;
    P[P_RO][p] = ro;
    P[P_PR][p] = pr;
    P[P_VX][p] = vx;
    P[P_VY][p] = vy;
    P[P_VZ][p] = vz;
    P[P_BX][p] = bx;
    P[P_BY][p] = by;
    P[P_BZ][p] = bz;
    return cfl ? cfl3_term(a.safety, a.gm, ro, pr, v2, b2, dg) : kR4_1e20;
}

// adddfdx.f for one cell: u <- u - dts*((F_hi - F_lo)/dd)
CANS_DEV void hlld_apply(double *const *u, size_t p, double dts, double dd, const double *hi, const double *lo) {
#pragma unroll
    for (int q = 0; q < 8; q++) u[q][p] = u[q][p] - dts * ((hi[q] - lo[q]) / dd);
}

// x sweep: a wave covers faces i = 63*strip + lane (lane 0 overlaps the
// previous strip), cells i = 2..ix-1 take their west face from lane-1.
// With FP64 fluxes it is compiled for four waves per SIMD (128 VGPRs, a few
// spilled): at 384^3 that took an FP64 step from 38.0 to 31.3 ms against the
// compiler's own choice (134 VGPRs, three waves).
template <class T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? 4 : 1))) void hlld_sweep_x(HlldArgs a, int stage,
                                                                                                          int cons) {
    double dt;
    // in a periodic run the stage-1 x sweep is the step's first kernel
    if (!hstep(a, dt, a.per_margin > 0 && stage == 1)) return;
    const double dts = stage == 1 ? 0.5 * dt : dt;   // engine.f: dth = 0.5d0*dt
    const int ix = a.ix, jx = a.jx, kx = a.kx;
    const int lane = threadIdx.x & 63;
    const int w = uni((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    const int nstrip = (ix + 62) / 63;
    if (w >= nstrip * jx * kx) return;                // whole waves
    const int strip = w % nstrip, row = w / nstrip;
    const int j = 1 + row % jx, k = 1 + row / jx;
    const int i = 63 * strip + lane;
    const size_t p = (size_t)(i - 1) + (size_t)ix * ((size_t)(j - 1) + (size_t)jx * (size_t)(k - 1));
    const double *P[8];
#pragma unroll
    for (int q = 0; q < 8; q++) P[q] = stage == 1 ? a.prim[q] : a.half[q];
    double F[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (i >= 1 && i <= ix - 1) hlld_face_flux<T>(a, P, 0, stage, i, ix, p, 1, F);
    double FW[8];
#pragma unroll
    for (int q = 0; q < 8; q++) FW[q] = shr1(F[q]);
    double *u[8];
#pragma unroll
    for (int q = 0; q < 8; q++) u[q] = stage == 1 ? a.u1[q] : a.u2[q];
    if (cons) {
        // the first sweep of the stage also starts u from the primitives
        // (hlld_cons fused): every cell i = 1..ix of the row
        if (lane >= 1 && i >= 1 && i <= ix) {
            double U[8];
            hlld_cons_of(a, p, U);
            if (i >= 2 && i <= ix - 1) {
                const double dd = a.dx[i - 1];
#pragma unroll
                for (int q = 0; q < 8; q++) u[q][p] = U[q] - dts * ((F[q] - FW[q]) / dd);
            } else {
#pragma unroll
                for (int q = 0; q < 8; q++) u[q][p] = U[q];
            }
        }
    } else if (lane >= 1 && i >= 2 && i <= ix - 1) {
        hlld_apply(u, p, dts, a.dx[i - 1], F, FW);
    }
}

// y or z sweep: thread (i, other transverse index, segment) marches over the
// cells cb..ce of its segment, computing faces cb-1..ce once each from a
// sliding window of four cells (8 loads per face instead of 32: at 384^3
// the FP64 step went from 31.3 to 26.4 ms, FP32-flux from 26.0 to 21.9 ms).
// The window needs two waves per SIMD with FP64 fluxes (231 VGPRs; at three
// the spills made it slower), three with FP32.
// The direction is a template argument: the face states and fluxes are
// indexed by it, and a run-time index would put those arrays in scratch.
template <class T, int DIR>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? 2 : 1))) void hlld_sweep_yz(HlldArgs a, int stage, int H,
                                                                                                           int to_prim) {
    constexpr int dir = DIR;
    __shared__ double red[4];
    // FP64 fluxes: the conserved state of the next cell goes one face ahead
    // into LDS by global->LDS loads (no registers held while the face is
    // computed); per wave 8 fields x (low, high word) x 64 lanes
    constexpr bool kLdsU = sizeof(T) == 8;
    __shared__ unsigned ubuf[kLdsU ? 4 : 1][16][64];
    double dt;
    if (!hstep(a, dt)) return;
    const double dts = stage == 1 ? 0.5 * dt : dt;
    const int ix = a.ix, jx = a.jx, kx = a.kx;
    const int nd = dir == 1 ? jx : kx, no = dir == 1 ? kx : jx;
    const int nseg = (nd - 2 + H - 1) / H;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    // the next step's cfl_m in a periodic run (the stage-2 z sweep, which
    // recovers the new primitives): every thread of the block reaches the
    // block minimum at the end
    const int mg = a.per_margin;
    const bool do_cfl = DIR == 2 && stage == 2 && to_prim && a.clk && mg > 0 && a.safety > 0.0;
    double dtmin_local = kR4_1e20;
    if (t < (size_t)ix * no * nseg) {
        const int i = 1 + (int)(t % ix);
        const int o = 1 + (int)((t / ix) % no);
        const int seg = (int)(t / ((size_t)ix * no));
        const int cb = 2 + seg * H, ce = min(cb + H - 1, nd - 1);
        const size_t s = dir == 1 ? (size_t)ix : (size_t)ix * jx;
        const size_t base = (size_t)(i - 1) + (dir == 1 ? (size_t)ix * jx * (size_t)(o - 1) : (size_t)ix * (size_t)(o - 1));
        const double *P[8];
        double *u[8];
#pragma unroll
        for (int q = 0; q < 8; q++) { P[q] = stage == 1 ? a.prim[q] : a.half[q]; u[q] = stage == 1 ? a.u1[q] : a.u2[q]; }
        const double *dd = dir == 1 ? a.dy : a.dz;
        // interior column [m+1, n-m]^2 of the cfl_m (periodic images: image_min)
        const bool cij = do_cfl && i >= mg + 1 && i <= ix - mg && o >= mg + 1 && o <= jx - mg;
        const double dxyg = cij ? fminF(image_min(a.dx, i, ix, mg), image_min(a.dy, o, jx, mg)) : 0.0;
        // cells c-1..c+2 of face c in a sliding window: one new cell per face
        // (cell indices clamped to [1, nd]; a clamped cell only feeds a face
        // whose states are piecewise constant, which does not read it)
        double w0[8], w1[8], w2[8], w3[8];
        auto load_cell = [&](int cell, double *d) {
            const int cc = cell < 1 ? 1 : (cell > nd ? nd : cell);
            const size_t pc = base + s * (size_t)(cc - 1);
#pragma unroll
            for (int q = 0; q < 8; q++) d[q] = P[q][pc];
        };
        // the next cell of the window is loaded one face ahead, so that its
        // loads are in flight while the current face is computed
        double wn[8];
        load_cell(cb - 2, w0); load_cell(cb - 1, w1); load_cell(cb, w2); load_cell(cb + 1, w3);
        load_cell(cb + 2, wn);
        double Fl[8], Fh[8];
        hlld_face_w<T>(a, dir, stage == 2 && cb - 1 >= 2 && cb - 1 <= nd - 2, w0, w1, w2, w3, Fl);
        // With FP32 fluxes (registers to spare) the conserved state of the next
        // cell is also loaded one face ahead (384^3 FP32-flux step 17.3 ->
        // 15.8 ms); with FP64 fluxes the extra registers spill (18.9 -> 29.7 ms).
        constexpr bool kPreU = sizeof(T) == 4;
        double un[8];
        if (kPreU) {
#pragma unroll
            for (int q = 0; q < 8; q++) un[q] = u[q][base + s * (size_t)(cb - 1)];
        }
        // each lane writes its own words of the wave's slots and reads only
        // them back; a read waits for every load of the wave in flight (the
        // compiler does not order LDS reads after global->LDS loads), by then
        // the window's next cell and the previous cell's stores have had the
        // whole face to complete
        const int wv = kLdsU ? (int)(threadIdx.x >> 6) : 0, lane = (int)(threadIdx.x & 63);
        auto pre_u = [&](size_t pn) {
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const char *g = (const char *)(u[q] + pn);
                __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)g,
                                                 (__attribute__((address_space(3))) void *)&ubuf[wv][2 * q][0], 4, 0, 0);
                __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(g + 4),
                                                 (__attribute__((address_space(3))) void *)&ubuf[wv][2 * q + 1][0], 4, 0, 0);
            }
        };
        auto get_u = [&](double *U0) {
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
#pragma unroll
            for (int q = 0; q < 8; q++)
                U0[q] = __builtin_bit_cast(double, ((unsigned long long)ubuf[wv][2 * q + 1][lane] << 32) |
                                                       (unsigned long long)ubuf[wv][2 * q][lane]);
        };
        if (kLdsU) pre_u(base + s * (size_t)(cb - 1));
        for (int c = cb; c <= ce; c++) {
#pragma unroll
            for (int q = 0; q < 8; q++) { w0[q] = w1[q]; w1[q] = w2[q]; w2[q] = w3[q]; w3[q] = wn[q]; }
            if (c < ce) load_cell(c + 3, wn);
            const size_t p = base + s * (size_t)(c - 1);
            if (kPreU) {
                double uc[8];
#pragma unroll
                for (int q = 0; q < 8; q++) uc[q] = un[q];
                if (c < ce) {
#pragma unroll
                    for (int q = 0; q < 8; q++) un[q] = u[q][p + s];
                }
                hlld_face_w<T>(a, dir, stage == 2 && c >= 2 && c <= nd - 2, w0, w1, w2, w3, Fh);
                double U[8];
                const double dl = dd[c - 1];
#pragma unroll
                for (int q = 0; q < 8; q++) U[q] = uc[q] - dts * ((Fh[q] - Fl[q]) / dl);   // adddfdx.f
                if (to_prim) {
                    const bool cc = cij && c >= mg + 1 && c <= nd - mg;
                    const double dq = hlld_prim_of(a, U, stage == 1 ? a.half : a.out, p, cc,
                                                   cc ? fminF(dxyg, image_min(dd, c, nd, mg)) : 0.0);
                    if (dq < dtmin_local) dtmin_local = dq;   // NaN never wins
                } else {
#pragma unroll
                    for (int q = 0; q < 8; q++) u[q][p] = U[q];
                }
#pragma unroll
                for (int q = 0; q < 8; q++) Fl[q] = Fh[q];
                continue;
            }
            // the spacings are loaded before the face is computed: the wait
            // of get_u would otherwise expose their latency
            const double dl = dd[c - 1];
            const bool cc = to_prim && cij && c >= mg + 1 && c <= nd - mg;
            const double dg = cc ? fminF(dxyg, image_min(dd, c, nd, mg)) : 0.0;
            hlld_face_w<T>(a, dir, stage == 2 && c >= 2 && c <= nd - 2, w0, w1, w2, w3, Fh);
            double U0[8];
            get_u(U0);
            if (to_prim) {
                // the last sweep of the stage also recovers the primitives
                // (hlld_prim fused): half (stage 1) or out (stage 2)
                double U[8];
#pragma unroll
                for (int q = 0; q < 8; q++) U[q] = U0[q] - dts * ((Fh[q] - Fl[q]) / dl);   // adddfdx.f
                if (c < ce) pre_u(p + s);
                const double dq = hlld_prim_of(a, U, stage == 1 ? a.half : a.out, p, cc, dg);
                if (dq < dtmin_local) dtmin_local = dq;   // NaN never wins
            } else {
#pragma unroll
                for (int q = 0; q < 8; q++) u[q][p] = U0[q] - dts * ((Fh[q] - Fl[q]) / dl);   // adddfdx.f
                if (c < ce) pre_u(p + s);
            }
#pragma unroll
            for (int q = 0; q < 8; q++) Fl[q] = Fh[q];
        }
        if (to_prim) {   // the end cells 1 and nd of the column, not swept
            for (int c = 1; c <= nd; c += nd - 1) {
                if ((c == 1 && cb != 2) || (c == nd && ce != nd - 1)) continue;
                const size_t p = base + s * (size_t)(c - 1);
                double U[8];
#pragma unroll
                for (int q = 0; q < 8; q++) U[q] = u[q][p];
                hlld_prim_of(a, U, stage == 1 ? a.half : a.out, p);
            }
        }
    }
    if (do_cfl) block_min_to_global(dtmin_local, red, &a.clk->dtkey[(a.step + 1) % 3]);
}

// rvtovv.f, ee2eh.f, eh2pr.f: primitives of every cell from u (stage 1 -> half,
// stage 2 -> prim)
__global__ __launch_bounds__(256) void hlld_prim(HlldArgs a, int stage) {
    double dt;
    if (!hstep(a, dt)) return;
    const size_t N = (size_t)a.ix * a.jx * a.kx;
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    double *const *u = stage == 1 ? a.u1 : a.u2;
    double U[8];
#pragma unroll
    for (int q = 0; q < 8; q++) U[q] = u[q][p];
    hlld_prim_of(a, U, stage == 1 ? a.half : a.out, p);
}

template <class T>
hipError_t launch_hlld_t(const HlldArgs &a, hipStream_t s) {
    const size_t N = (size_t)a.ix * a.jx * a.kx;
    const int n[3] = {a.ix, a.jx, a.kx};
    const unsigned gN = (unsigned)((N + 255) / 256);
    // u1, u2 = U(prim): fused into the x sweeps when x is the first sweep;
    // the primitives of each stage: fused into the z sweeps when z is the last
    const int fuse = a.mfdim[0] == 1 && n[0] >= 3;
    const int fuse_p = a.mfdim[2] == 1 && n[2] >= 3;
    // a periodic run's clock and cfl_m ride on the x sweep and the z sweep
    if (a.per_margin > 0 && !(fuse && fuse_p)) return hipErrorInvalidValue;
    if (!fuse) hipLaunchKernelGGL(hlld_cons, dim3(gN), dim3(256), 0, s, a);
    for (int stage = 1; stage <= 2; stage++) {
        for (int dir = 0; dir < 3; dir++) {
            if (a.mfdim[dir] != 1) continue;
            if (n[dir] < 3) continue;                 // no interior cell along dir
            if (dir == 0) {
                const size_t waves = (size_t)((a.ix + 62) / 63) * a.jx * a.kx;
                hipLaunchKernelGGL(hlld_sweep_x<T>, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, a, stage, fuse);
            } else {
                // segments so that the launch has >= march_waves() waves (a
                // sweep of 2048, 8192, 32768, 131072 at 384^3: 32768 is fastest)
                const size_t cols = N / n[dir];
                int nseg = (int)(((unsigned long long)march_waves() * 64 + cols - 1) / cols);
                nseg = nseg < 1 ? 1 : (nseg > n[dir] - 2 ? n[dir] - 2 : nseg);
                const int H = (n[dir] - 2 + nseg - 1) / nseg;
                nseg = (n[dir] - 2 + H - 1) / H;
                const size_t nt = cols * (size_t)nseg;
                if (dir == 1)
                    hipLaunchKernelGGL((hlld_sweep_yz<T, 1>), dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, a,
                                       stage, H, 0);
                else
                    hipLaunchKernelGGL((hlld_sweep_yz<T, 2>), dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, a,
                                       stage, H, fuse_p);
            }
        }
        if (!fuse_p) hipLaunchKernelGGL(hlld_prim, dim3(gN), dim3(256), 0, s, a, stage);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_hlld(const HlldArgs &a, bool mixed, hipStream_t s) {
    return mixed ? launch_hlld_t<float>(a, s) : launch_hlld_t<double>(a, s);
}

}  // namespace cans

// ======================================================================
// hlld_flux and hlld_mh with their own reference argument lists, for callers
// that compose an engine from them (the fused engine above does not use
// these).  One thread per array element.
namespace cans {
int abi_fail(int code, const char *fmt, ...);
int abi_scratch(size_t bytes, void **out, hipStream_t s);
namespace {

// hdhlld/hlld_flux.f90:141-143: every face i,j,k; w arrays hold the left
// state, then the right one, n apart
__global__ void hlld_flux_k(const double *row, const double *prw, const double *vxw, const double *vyw,
                            const double *vzw, const double *bx, const double *byw, const double *bzw, double gm,
                            double zfloor, double *fro, double *fee, double *frx, double *fry, double *frz,
                            double *fby, double *fbz, size_t n) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const size_t q = p + n;
    double o[7];
    hlld_face<double>(gm, zfloor, bx[p], row[p], prw[p], vxw[p], vyw[p], vzw[p], byw[p], bzw[p], row[q], prw[q],
                      vxw[q], vyw[q], vzw[q], byw[q], bzw[q], o);
    fro[p] = o[0]; fee[p] = o[1]; frx[p] = o[2]; fry[p] = o[3]; frz[p] = o[4]; fby[p] = o[5]; fbz[p] = o[6];
}

__global__ void nan_fill_k(double *p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = __builtin_nan("");
}

struct MhDir {
    const double *prim[8];   // rof..bzf
    double *un[8];           // ron..bzn
    double *w0[8], *w1[8];   // interpolated left/right states
    double *bb;              // face-normal field
    double *f[8];            // fluxes (ro ee rx ry rz bx by bz)
    const double *dd;        // dx, dy or dz
    double gm, dt;
    int ix, jx, kx, mdir, mstage;
};

CANS_DEV int along(const MhDir &a, int i, int j, int k) { return a.mdir == 1 ? i : (a.mdir == 2 ? j : k); }
CANS_DEV size_t stride(const MhDir &a) {
    return a.mdir == 1 ? 1 : (a.mdir == 2 ? (size_t)a.ix : (size_t)a.ix * a.jx);
}

// interpolation/inp_const.f (mstage 1) and inp_tvdminmod.f (mstage 2, with the
// constant faces at 1 and n-1), then the face-normal field of hlld_mh.f:43-95
__global__ void mh_interp_k(MhDir a) {
    const size_t n = (size_t)a.ix * a.jx * a.kx;
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int i = (int)(p % (size_t)a.ix) + 1, j = (int)((p / (size_t)a.ix) % (size_t)a.jx) + 1;
    const int k = (int)(p / ((size_t)a.ix * a.jx)) + 1;
    const int x = along(a, i, j, k), nd = a.mdir == 1 ? a.ix : (a.mdir == 2 ? a.jx : a.kx);
    if (x > nd - 1) return;
    const size_t s = stride(a);
    for (int q = 0; q < 8; q++) {
        const double *da = a.prim[q];
        if (a.mstage == 1 || x == 1 || x == nd - 1) {
            a.w0[q][p] = da[p];
            a.w1[q][p] = da[p + s];
        } else {
            a.w0[q][p] = da[p] + 0.5 * flmtH(da[p] - da[p - s], da[p + s] - da[p]);
            a.w1[q][p] = da[p + s] - 0.5 * flmtH(da[p + 2 * s] - da[p + s], da[p + s] - da[p]);
        }
    }
    const double *bn = a.prim[a.mdir == 1 ? 5 : (a.mdir == 2 ? 6 : 7)];
    a.bb[p] = (bn[p] + bn[p + s]) / 2.0;
}

// hlld_flux over every face in the sweep frame, then the normal-field flux
// zeroed on i = 1..ix-1 (hlld_mh.f:56-94)
__global__ void mh_flux_k(MhDir a) {
    const size_t n = (size_t)a.ix * a.jx * a.kx;
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int vn, v1, v2, b1, b2, fn, f1, f2, g1, g2, gn;
    if (a.mdir == 1) { vn = 2; v1 = 3; v2 = 4; b1 = 6; b2 = 7; fn = 2; f1 = 3; f2 = 4; g1 = 6; g2 = 7; gn = 5; }
    else if (a.mdir == 2) { vn = 3; v1 = 4; v2 = 2; b1 = 7; b2 = 5; fn = 3; f1 = 4; f2 = 2; g1 = 7; g2 = 5; gn = 6; }
    else { vn = 4; v1 = 2; v2 = 3; b1 = 5; b2 = 6; fn = 4; f1 = 2; f2 = 3; g1 = 5; g2 = 6; gn = 7; }
    double o[7];
    hlld_face<double>(a.gm, 1.0e-10, a.bb[p], a.w0[0][p], a.w0[1][p], a.w0[vn][p], a.w0[v1][p], a.w0[v2][p],
                      a.w0[b1][p], a.w0[b2][p], a.w1[0][p], a.w1[1][p], a.w1[vn][p], a.w1[v1][p], a.w1[v2][p],
                      a.w1[b1][p], a.w1[b2][p], o);
    a.f[0][p] = o[0]; a.f[1][p] = o[1]; a.f[fn][p] = o[2]; a.f[f1][p] = o[3]; a.f[f2][p] = o[4];
    a.f[g1][p] = o[5]; a.f[g2][p] = o[6];
    const int i = (int)(p % (size_t)a.ix) + 1;
    if (i <= a.ix - 1) a.f[gn][p] = 0.0;
}

// common/adddfdx.f
__global__ void mh_add_k(MhDir a) {
    const size_t n = (size_t)a.ix * a.jx * a.kx;
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int i = (int)(p % (size_t)a.ix) + 1, j = (int)((p / (size_t)a.ix) % (size_t)a.jx) + 1;
    const int k = (int)(p / ((size_t)a.ix * a.jx)) + 1;
    const int x = along(a, i, j, k), nd = a.mdir == 1 ? a.ix : (a.mdir == 2 ? a.jx : a.kx);
    if (x < 2 || x > nd - 1) return;
    const size_t s = stride(a);
    const double d = a.dd[x - 1];
    for (int q = 0; q < 8; q++) a.un[q][p] = a.un[q][p] - a.dt * ((a.f[q][p] - a.f[q][p - s]) / d);
}

}  // namespace
}  // namespace cans

extern "C" {

int cans3d_hlld_flux(const double *row, const double *prw, const double *vxw, const double *vyw, const double *vzw,
                     const double *bx, const double *byw, const double *bzw, double gm, int ix, int jx, int kx,
                     double zfloor, double *fro, double *fee, double *frx, double *fry, double *frz, double *fby,
                     double *fbz, void *stream) {
    using namespace cans;
    if (ix < 1 || jx < 1 || kx < 1) return abi_fail(CANS_ERR_ARG, "cans3d_hlld_flux: ix=%d jx=%d kx=%d", ix, jx, kx);
    const size_t n = (size_t)ix * jx * kx;
    hipLaunchKernelGGL(hlld_flux_k, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, row, prw,
                       vxw, vyw, vzw, bx, byw, bzw, gm, zfloor, fro, fee, frx, fry, frz, fby, fbz, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return abi_fail(CANS_ERR_HIP, "cans3d_hlld_flux: %s", hipGetErrorString(e));
    return 0;
}

int cans3d_hlld_mh(const double *rof, const double *prf, const double *vxf, const double *vyf, const double *vzf,
                   const double *bxf, const double *byf, const double *bzf, double *ron, double *een, double *rxn,
                   double *ryn, double *rzn, double *bxn, double *byn, double *bzn, double gm, double dt,
                   const double *dx, const double *dy, const double *dz, int ix, int jx, int kx, const int *mfdim,
                   int mstage, void *stream) {
    using namespace cans;
    if (ix < 1 || jx < 1 || kx < 1 || !mfdim || (mstage != 1 && mstage != 2))
        return abi_fail(CANS_ERR_ARG, "cans3d_hlld_mh: ix=%d jx=%d kx=%d mstage=%d", ix, jx, kx, mstage);
    hipStream_t s = (hipStream_t)stream;
    const size_t n = (size_t)ix * jx * kx;
    // the reference's automatic arrays (w0, w1, bb, f), NaN where it leaves
    // them unset; kept across the directions of the call as the reference's are
    void *w = nullptr;
    if (int rc = abi_scratch(25 * n * sizeof(double), &w, s)) return rc;
    double *ws = (double *)w;
    hipLaunchKernelGGL(nan_fill_k, dim3(1024), dim3(256), 0, s, ws, 25 * n);
    MhDir a{};
    const double *prim[8] = {rof, prf, vxf, vyf, vzf, bxf, byf, bzf};
    double *un[8] = {ron, een, rxn, ryn, rzn, bxn, byn, bzn};
    for (int q = 0; q < 8; q++) {
        a.prim[q] = prim[q];
        a.un[q] = un[q];
        a.w0[q] = ws + q * n;
        a.w1[q] = ws + (8 + q) * n;
        a.f[q] = ws + (17 + q) * n;
    }
    a.bb = ws + 16 * n;
    a.gm = gm; a.dt = dt; a.ix = ix; a.jx = jx; a.kx = kx; a.mstage = mstage;
    const dim3 g((unsigned)((n + 255) / 256));
    for (int d = 1; d <= 3; d++) {
        if (mfdim[d - 1] != 1) continue;
        a.mdir = d;
        a.dd = d == 1 ? dx : (d == 2 ? dy : dz);
        hipLaunchKernelGGL(mh_interp_k, g, dim3(256), 0, s, a);
        hipLaunchKernelGGL(mh_flux_k, g, dim3(256), 0, s, a);
        hipLaunchKernelGGL(mh_add_k, g, dim3(256), 0, s, a);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return abi_fail(CANS_ERR_HIP, "cans3d_hlld_mh: %s", hipGetErrorString(e));
    return 0;
}

}  // extern "C"
