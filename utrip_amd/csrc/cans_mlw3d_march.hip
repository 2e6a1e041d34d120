// cans_mlw3d_march.hip -- the Cartesian 3D modified Lax-Wendroff step
// (cans3d/hdmlw/mlw_m.f with mlwhalf.f:38-68, mlwfull.f:34-55,
// mlwartv.f:36-48) as one kernel that marches in z, bit-exact FP64.
//
// A block of 16x16 threads owns a 16x16 column tile of cells (a one-cell
// halo around 14x14 updated cells) and walks it through a z-segment of the
// grid.  Step s:
//   P1  every thread turns the primitives of its cell in plane s into the 23
//       cell values (conserved variables, E = -v x B, the cell fluxes) and
//       puts them in LDS; the primitives of plane s+1 are loaded meanwhile;
//   P2  corner s-1 (between cell planes s-1 and s) of each of the 15x15
//       corners from the eight cells around it: the half step (mlwhalf), its
//       primitives and the 18 corner fluxes, into LDS;
//   P3  update of the cell of plane s-1 from corner planes s-2 and s-1, its
//       primitives to HBM; then the parts of the update of the cell of plane
//       s that need only plane s (centre x and y terms, x and y viscosity).
// Cell planes are kept in two LDS buffers by parity (the corner reads both
// planes, and P3 reads plane s while P1 of the next step fills the other);
// the corner plane in one, except its z fluxes (the next step's lower plane)
// in two.  What the update of a cell needs from the corner plane below it
// (the x and y corner differences, combined with the cell's metric) is
// carried in registers from the step before, as are the parts of its own
// plane.  Every sum keeps the association of mlw_m.f: a left-to-right sum
// over the lower plane and then the upper one is continued where it stops,
// so each value is rounded as in the reference.
// Each segment starts two planes before its first updated cell; the host
// picks the number of z-segments so that the launch fills the CUs in the
// fewest plane-steps (one block per CU: its LDS is 152 KB).
// In a periodic box (per_margin = m > 0) the kernel advances the device
// clock itself and takes the next step's cfl_m (cans3d/common/cfl_m.f:39-58)
// over the cells it updates, as the 2D kernel does (image_min, cfl3_term).
#include "cans_common.hpp"
#include "cans_clock.hpp"
#include "cans_roe.hpp"

namespace cans {
namespace {

constexpr double pi8i = 1.0 / kPi / 8.0;
constexpr double pi4i = 1.0 / kPi / 4.0;
constexpr int TB = 16;          // tile edge in cells, halo included
constexpr int TO = TB - 2;      // updated cells per tile edge
constexpr int NT = TB * TB;     // threads per block

// LDS slots of a cell plane (two buffers, by plane parity)
enum { L_RO, L_EE, L_RX, L_RY, L_BX, L_BY, L_BZ, L_EEX, L_EEY, L_RXX, L_RXY, L_RYX, L_RYY, L_RZX, L_RZY, L_EZ,
       L_VX, L_VY, L_NC };
// z-part of a cell plane (two buffers): rz eez rxz ryz rzz ey ex
enum { M_RZ, M_EEZ, M_RXZ, M_RYZ, M_RZZ, M_EY, M_EX, M_N };
// x and y parts of a corner plane (one buffer)
enum { K_ROX, K_EEX, K_RXX, K_RYX, K_RZX, K_ROY, K_EEY, K_RXY, K_RYY, K_RZY, K_EZ, K_N };
// z-part of a corner plane (two buffers): roz eez rxz ryz rzz ey ex, slots M_*

struct CellV {
    double u[8];   // ro ee rx ry rz bx by bz
    double eex, eey, eez, rxx, rxy, rxz, ryx, ryy, ryz, rzx, rzy, rzz, ex, ey, ez, vx, vy, vz;
};
struct CornV {
    double rox, roy, roz, eex, eey, eez, rxx, rxy, rxz, ryx, ryy, ryz, rzx, rzy, rzz, ex, ey, ez;
};

// cell values from primitives (the cell loops of mlw_m.f:70-192)
CANS_DEV void cell_values(const double *w, double gm, CellV &c) {
    const double ro = w[0], pr = w[1], vx = w[2], vy = w[3], vz = w[4], bx = w[5], by = w[6], bz = w[7];
    const double vv = vx * vx + vy * vy + vz * vz;
    const double bb = bx * bx + by * by + bz * bz;
    c.u[0] = ro;
    c.u[1] = pr / (gm - 1.0) + 0.5 * ro * vv + pi8i * bb;
    c.u[2] = ro * vx;
    c.u[3] = ro * vy;
    c.u[4] = ro * vz;
    c.u[5] = bx;
    c.u[6] = by;
    c.u[7] = bz;
    const double ex = -vy * bz + vz * by;
    const double ey = -vz * bx + vx * bz;
    const double ez = -vx * by + vy * bx;
    const double ep = pr * gm / (gm - 1.0) + 0.5 * ro * vv;
    c.eex = ep * vx + (bz * ey - by * ez) * pi4i;
    c.eey = ep * vy + (bx * ez - bz * ex) * pi4i;
    c.eez = ep * vz + (by * ex - bx * ey) * pi4i;
    c.rxx = ro * (vx * vx) + pr + pi8i * (by * by + bz * bz - bx * bx);
    c.rxy = ro * vx * vy - pi4i * bx * by;
    c.rxz = ro * vx * vz - pi4i * bx * bz;
    c.ryx = ro * vy * vx - pi4i * by * bx;
    c.ryy = ro * (vy * vy) + pr + pi8i * (bx * bx + bz * bz - by * by);
    c.ryz = ro * vy * vz - pi4i * by * bz;
    c.rzx = ro * vz * vx - pi4i * bz * bx;
    c.rzy = ro * vz * vy - pi4i * bz * by;
    c.rzz = ro * (vz * vz) + pr + pi8i * (bx * bx + by * by - bz * bz);
    c.ex = ex;
    c.ey = ey;
    c.ez = ez;
    c.vx = vx;
    c.vy = vy;
    c.vz = vz;
}

struct Lds {
    double c[2 * L_NC * NT];    // cell planes, by parity
    double cz[2 * M_N * NT];    // z-part of cell planes, by parity
    double h[K_N * NT];         // corner plane
    double hz[2 * M_N * NT];    // z-part of corner planes, by parity
};

CANS_DEV void put_cell(Lds &L, int t, int par, const CellV &c) {
    double *p = L.c + par * L_NC * NT + t;
    p[L_RO * NT] = c.u[0];
    p[L_EE * NT] = c.u[1];
    p[L_RX * NT] = c.u[2];
    p[L_RY * NT] = c.u[3];
    p[L_BX * NT] = c.u[5];
    p[L_BY * NT] = c.u[6];
    p[L_BZ * NT] = c.u[7];
    p[L_EEX * NT] = c.eex;
    p[L_EEY * NT] = c.eey;
    p[L_RXX * NT] = c.rxx;
    p[L_RXY * NT] = c.rxy;
    p[L_RYX * NT] = c.ryx;
    p[L_RYY * NT] = c.ryy;
    p[L_RZX * NT] = c.rzx;
    p[L_RZY * NT] = c.rzy;
    p[L_EZ * NT] = c.ez;
    p[L_VX * NT] = c.vx;
    p[L_VY * NT] = c.vy;
    double *z = L.cz + par * M_N * NT + t;
    z[M_RZ * NT] = c.u[4];
    z[M_EEZ * NT] = c.eez;
    z[M_RXZ * NT] = c.rxz;
    z[M_RYZ * NT] = c.ryz;
    z[M_RZZ * NT] = c.rzz;
    z[M_EY * NT] = c.ey;
    z[M_EX * NT] = c.ex;
}
// conserved variable q of the cell at t in the plane of parity par
CANS_DEV double ldu(const Lds &L, int t, int par, int q) {
    const double *p = L.c + par * L_NC * NT + t;
    switch (q) {
    case 0: return p[L_RO * NT];
    case 1: return p[L_EE * NT];
    case 2: return p[L_RX * NT];
    case 3: return p[L_RY * NT];
    case 4: return L.cz[(par * M_N + M_RZ) * NT + t];
    case 5: return p[L_BX * NT];
    case 6: return p[L_BY * NT];
    default: return p[L_BZ * NT];
    }
}
// signed flux of equation q in direction d of that cell (0.0 where mlw_m.f
// has none)
CANS_DEV double ldf(const Lds &L, int t, int par, int q, int d) {
    const double *p = L.c + par * L_NC * NT + t;
    const double *z = L.cz + par * M_N * NT + t;
    switch (q * 3 + d) {
    case 0: return p[L_RX * NT];
    case 1: return p[L_RY * NT];
    case 2: return z[M_RZ * NT];
    case 3: return p[L_EEX * NT];
    case 4: return p[L_EEY * NT];
    case 5: return z[M_EEZ * NT];
    case 6: return p[L_RXX * NT];
    case 7: return p[L_RXY * NT];
    case 8: return z[M_RXZ * NT];
    case 9: return p[L_RYX * NT];
    case 10: return p[L_RYY * NT];
    case 11: return z[M_RYZ * NT];
    case 12: return p[L_RZX * NT];
    case 13: return p[L_RZY * NT];
    case 14: return z[M_RZZ * NT];
    case 16: return p[L_EZ * NT];
    case 17: return -z[M_EY * NT];
    case 18: return -p[L_EZ * NT];
    case 20: return z[M_EX * NT];
    case 21: return z[M_EY * NT];
    case 22: return -z[M_EX * NT];
    default: return 0.0;
    }
}
CANS_DEV double ldv(const Lds &L, int t, int par, int slot) { return L.c[(par * L_NC + slot) * NT + t]; }
CANS_DEV void put_corner(Lds &L, int t, int par, const CornV &h) {
    double *p = L.h + t;
    p[K_ROX * NT] = h.rox;
    p[K_EEX * NT] = h.eex;
    p[K_RXX * NT] = h.rxx;
    p[K_RYX * NT] = h.ryx;
    p[K_RZX * NT] = h.rzx;
    p[K_ROY * NT] = h.roy;
    p[K_EEY * NT] = h.eey;
    p[K_RXY * NT] = h.rxy;
    p[K_RYY * NT] = h.ryy;
    p[K_RZY * NT] = h.rzy;
    p[K_EZ * NT] = h.ez;
    double *z = L.hz + par * M_N * NT + t;
    z[M_RZ * NT] = h.roz;
    z[M_EEZ * NT] = h.eez;
    z[M_RXZ * NT] = h.rxz;
    z[M_RYZ * NT] = h.ryz;
    z[M_RZZ * NT] = h.rzz;
    z[M_EY * NT] = h.ey;
    z[M_EX * NT] = h.ex;
}
// signed flux of equation q in direction d of the corner at t, z-part from
// the corner plane of parity par
CANS_DEV double ldh(const Lds &L, int t, int par, int q, int d) {
    const double *p = L.h + t;
    const double *z = L.hz + par * M_N * NT + t;
    switch (q * 3 + d) {
    case 0: return p[K_ROX * NT];
    case 1: return p[K_ROY * NT];
    case 2: return z[M_RZ * NT];
    case 3: return p[K_EEX * NT];
    case 4: return p[K_EEY * NT];
    case 5: return z[M_EEZ * NT];
    case 6: return p[K_RXX * NT];
    case 7: return p[K_RXY * NT];
    case 8: return z[M_RXZ * NT];
    case 9: return p[K_RYX * NT];
    case 10: return p[K_RYY * NT];
    case 11: return z[M_RYZ * NT];
    case 12: return p[K_RZX * NT];
    case 13: return p[K_RZY * NT];
    case 14: return z[M_RZZ * NT];
    case 16: return p[K_EZ * NT];
    case 17: return -z[M_EY * NT];
    case 18: return -p[K_EZ * NT];
    case 20: return z[M_EX * NT];
    case 21: return z[M_EY * NT];
    case 22: return -z[M_EX * NT];
    default: return 0.0;
    }
}

// corner primitives and fluxes from the half-step conserved variables
// (mlw_m.f:193-315)
CANS_DEV void corner_fluxes(const double *un, double gm, CornV &h) {
    const double roh = un[0], eeh = un[1], bxh = un[5], byh = un[6], bzh = un[7];
    const double vxh = un[2] / roh;
    const double vyh = un[3] / roh;
    const double vzh = un[4] / roh;
    const double vv = vxh * vxh + vyh * vyh + vzh * vzh;
    const double bb = bxh * bxh + byh * byh + bzh * bzh;
    const double prh = (gm - 1.0) * (eeh - 0.5 * roh * vv - pi8i * bb);
    const double exh = -vyh * bzh + vzh * byh;
    const double eyh = -vzh * bxh + vxh * bzh;
    const double ezh = -vxh * byh + vyh * bxh;
    const double ep = prh * gm / (gm - 1.0) + 0.5 * roh * vv;
    h.rox = roh * vxh;
    h.roy = roh * vyh;
    h.roz = roh * vzh;
    h.eex = ep * vxh + (bzh * eyh - byh * ezh) * pi4i;
    h.eey = ep * vyh + (bxh * ezh - bzh * exh) * pi4i;
    h.eez = ep * vzh + (byh * exh - bxh * eyh) * pi4i;
    h.rxx = roh * (vxh * vxh) + prh + pi8i * (byh * byh + bzh * bzh - bxh * bxh);
    h.rxy = roh * vxh * vyh - pi4i * bxh * byh;
    h.rxz = roh * vxh * vzh - pi4i * bxh * bzh;
    h.ryx = roh * vyh * vxh - pi4i * byh * bxh;
    h.ryy = roh * (vyh * vyh) + prh + pi8i * (bxh * bxh + bzh * bzh - byh * byh);
    h.ryz = roh * vyh * vzh - pi4i * byh * bzh;
    h.rzx = roh * vzh * vxh - pi4i * bzh * bxh;
    h.rzy = roh * vzh * vyh - pi4i * bzh * byh;
    h.rzz = roh * (vzh * vzh) + prh + pi8i * (bxh * bxh + byh * byh - bzh * bzh);
    h.ex = exh;
    h.ey = eyh;
    h.ez = ezh;
}

// Step n's run decision and dt: the fixed dt without a clock; in a periodic
// box the clock is advanced here (block 0 writes it); otherwise a clock
// kernel of the step has advanced it.
CANS_DEV bool march_dt(const Mlw3dArgs &a, double &dt) {
    if (!a.clk) { dt = a.dt; return true; }
    const StepCtl c = clock_use(a.clk, a.step, a.per_margin == 0, blockIdx.x == 0 && threadIdx.x == 0);
    dt = c.dt;
    return c.run;
}


// One block: the column tile `tile` over the z-segment [k0, k1) of updated
// cells.  The march starts at cell plane k0-1: its first two steps form the
// corner plane k0-1 and the own-plane parts of cell k0, which the first
// update (at the step of plane k0+1) needs.
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(1, 1))) void mlw3d_march(Mlw3dArgs a,
                                                                                               int tiles_x,
                                                                                               int ntiles,
                                                                                               int nseg,
                                                                                               int seglen) {
    const unsigned per = gridDim.x / kXcds;
    const int item = (int)((blockIdx.x % kXcds) * per + blockIdx.x / kXcds);
    if (item >= ntiles * nseg) return;
    double dt;
    if (!march_dt(a, dt)) return;   // uniform across the launch
    __shared__ Lds L;
    __shared__ double red[NT / 64];
    const int ix = a.ix, jx = a.jx, kx = a.kx;
    const int tile = item % ntiles;
    const int k0 = 1 + (item / ntiles) * seglen;
    const int k1 = min(kx - 1, k0 + seglen);
    const double gm = a.gm, qav = a.qav, zero = 0.0;
    const int t = threadIdx.x, ti = t % TB, tj = t / TB;
    const int gx = (tile % tiles_x) * TO + ti;   // 0-based cell of this thread
    const int gy = (tile / tiles_x) * TO + tj;
    const bool cval = gx < ix && gy < jx;
    // corner (gx+1/2, gy+1/2): cells gx..gx+1, gy..gy+1
    const bool kval = ti < TB - 1 && tj < TB - 1 && gx <= ix - 2 && gy <= jx - 2;
    // updated cell: 1 <= gx, gy <= n-2, corners gx-1..gx (uval implies kval)
    const bool uval = ti >= 1 && tj >= 1 && ti <= TO && tj <= TO && gx <= ix - 2 && gy <= jx - 2;
    const size_t sy = ix, sz = (size_t)ix * jx;
    const size_t p0 = (size_t)gx + sy * gy;
    // the column's metric
    double dxim = 0, dyim = 0, dxi = 0, dyi = 0, dxim1 = 0, dyim1 = 0, ux0 = 0, ux1 = 0, uy0 = 0, uy1 = 0;
    double qx0 = 0, qx1 = 0, qy0 = 0, qy1 = 0;
    if (kval) { dxim = a.mx.dim[gx]; dyim = a.my.dim[gy]; }
    if (uval) {
        dxi = a.mx.di[gx]; dyi = a.my.di[gy];
        dxim1 = a.mx.dim[gx - 1]; dyim1 = a.my.dim[gy - 1];
        ux0 = a.mx.u0[gx]; ux1 = a.mx.u1[gx]; uy0 = a.my.u0[gy]; uy1 = a.my.u1[gy];
        qx0 = a.dxm[gx]; qx1 = a.dxm[gx - 1]; qy0 = a.dym[gy]; qy1 = a.dym[gy - 1];
    }
    // the next step's cfl_m: interior cells [m+1, n-m] (1-based) only
    const int mg = a.per_margin;
    const bool do_cfl = a.clk && mg > 0 && a.safety > 0.0;
    const bool cxy = do_cfl && uval && gx >= mg && gx < ix - mg && gy >= mg && gy < jx - mg;
    double dxyg = 0.0, dtmin_local = kR4_1e20;
    if (cxy) dxyg = fminF(image_min(a.dx, gx + 1, ix, mg), image_min(a.dy, gy + 1, jx, mg));
    // carried from step to step
    double dxy[8], avx[8], avy[8], bzl[8];       // cell s: parts from its own plane
    double fz2[8], alo[8], d0[8], d1[8];         // own z fluxes of the plane below; lower corner-plane parts
    double vzp = 0.0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        dxy[q] = avx[q] = avy[q] = bzl[q] = 0.0;
        fz2[q] = alo[q] = d0[q] = d1[q] = 0.0;
    }
    const int s0 = k0 - 1;
    double nxt[8];
    if (cval) {
#pragma unroll
        for (int f = 0; f < 8; f++) nxt[f] = a.prim[f][p0 + sz * s0];
    }
    for (int s = s0; s <= k1; s++) {
        const int par = s & 1, lp = par ^ 1;
        // metric of plane s-1 (corner s-1, cell s-1 and the own-plane parts of cell s)
        double zdim = 0.0, zdi = 0.0, zu0 = 0.0, zu1 = 0.0, zq = 0.0;
        if (s > s0) { zdim = a.mz.dim[s - 1]; zdi = a.mz.di[s - 1]; zu0 = a.mz.u0[s - 1]; zu1 = a.mz.u1[s - 1]; zq = a.dzm[s - 1]; }
        // P1: cell values of plane s
        double vzs = 0.0;   // vz of this thread's cell in plane s
        if (cval) {
            CellV cur;
            cell_values(nxt, gm, cur);
            vzs = nxt[4];
            if (s < k1) {
                const size_t pn = p0 + sz * (s + 1);
#pragma unroll
                for (int f = 0; f < 8; f++) nxt[f] = a.prim[f][pn];
            }
            put_cell(L, t, par, cur);
        }
        __syncthreads();
        // P2: corner s-1 from its eight cells in planes s-1 and s (each sum
        // over the lower plane first, continued over the upper one)
        if (kval) {
            const int t1 = t + 1, t2 = t + TB, t3 = t + TB + 1;   // cells 10, 01, 11 of the corner
            if (s > s0) {
                double un[8];
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const double uhp = ((ldu(L, t1, lp, q) + ldu(L, t, lp, q)) + ldu(L, t3, lp, q)) + ldu(L, t2, lp, q);
                    const double fxp = ((ldf(L, t1, lp, q, 0) - ldf(L, t, lp, q, 0)) + ldf(L, t3, lp, q, 0))
                                       - ldf(L, t2, lp, q, 0);
                    const double fyp = ((ldf(L, t2, lp, q, 1) - ldf(L, t, lp, q, 1)) + ldf(L, t3, lp, q, 1))
                                       - ldf(L, t1, lp, q, 1);
                    const double uh = ((((uhp + ldu(L, t1, par, q)) + ldu(L, t, par, q)) + ldu(L, t3, par, q))
                                       + ldu(L, t2, par, q)) / 8.0;
                    const double dfdx = ((((fxp + ldf(L, t1, par, q, 0)) - ldf(L, t, par, q, 0))
                                          + ldf(L, t3, par, q, 0)) - ldf(L, t2, par, q, 0)) / 4.0 * dxim;
                    const double dfdy = ((((fyp + ldf(L, t2, par, q, 1)) - ldf(L, t, par, q, 1))
                                          + ldf(L, t3, par, q, 1)) - ldf(L, t1, par, q, 1)) / 4.0 * dyim;
                    const double dfdz = (ldf(L, t, par, q, 2) - ldf(L, t, lp, q, 2) + ldf(L, t1, par, q, 2)
                                         - ldf(L, t1, lp, q, 2) + ldf(L, t2, par, q, 2) - ldf(L, t2, lp, q, 2)
                                         + ldf(L, t3, par, q, 2) - ldf(L, t3, lp, q, 2)) / 4.0 * zdim;
                    un[q] = uh - dt * dfdx - dt * dfdy - dt * dfdz;
                }
                CornV h;
                corner_fluxes(un, gm, h);
                put_corner(L, t, lp, h);
            }
        }
        __syncthreads();
        // P3: update of cell s-1 from corner planes s-2 (lower) and s-1
        // (upper); then the parts of the update of cell s that only need
        // plane s.  Neighbours' entries read here are not overwritten before
        // the next step's first barrier (P1 of step s+1 writes the other cell
        // buffer); the own column's entries of both cell buffers are written
        // only by this thread.
        if (uval) {
            const int c00 = t - TB - 1, c10 = t - TB, c01 = t - 1, c11 = t;
            if (s >= k0) {
                const bool upd = s > k0;   // cell s-1 in [k0, k1)
                const double qzc = qav * zq * fmaxF(zero, __builtin_fabs(vzs - vzp) - kR4_1em4);
                double U[8];
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const double aup = uy0 * (ldh(L, c10, lp, q, 0) - ldh(L, c00, lp, q, 0))
                                       + uy1 * (ldh(L, c11, lp, q, 0) - ldh(L, c01, lp, q, 0));
                    const double e0 = ldh(L, c01, lp, q, 1) - ldh(L, c00, lp, q, 1);
                    const double e1 = ldh(L, c11, lp, q, 1) - ldh(L, c10, lp, q, 1);
                    if (upd) {
                        const double ucq = ldu(L, t, lp, q);
                        double du = dxy[q] - 0.5 * dt * (0.5 * zdi * (ldf(L, t, par, q, 2) - fz2[q]));
                        const double dfdx = dxi * (zu0 * alo[q] + zu1 * aup);
                        const double dfdy = dyi * (ux0 * (zu0 * d0[q] + zu1 * e0) + ux1 * (zu0 * d1[q] + zu1 * e1));
                        const double dfdz =
                            zdi * (uy0 * (ux0 * (ldh(L, c00, lp, q, 2) - ldh(L, c00, par, q, 2))
                                          + ux1 * (ldh(L, c10, lp, q, 2) - ldh(L, c10, par, q, 2)))
                                   + uy1 * (ux0 * (ldh(L, c01, lp, q, 2) - ldh(L, c01, par, q, 2))
                                            + ux1 * (ldh(L, c11, lp, q, 2) - ldh(L, c11, par, q, 2))));
                        du = du - 0.5 * dt * dfdx - 0.5 * dt * dfdy - 0.5 * dt * dfdz;
                        du = du + avx[q] + avy[q] + dt * (zdi * (qzc * zdim * (ldu(L, t, par, q) - ucq) - bzl[q]));
                        U[q] = ucq + du;
                    }
                    alo[q] = aup;
                    d0[q] = e0;
                    d1[q] = e1;
                }
                if (upd) {
                    const size_t pk = p0 + sz * (s - 1);
                    const double ro = U[0];
                    const double vx = U[2] / ro;
                    const double vy = U[3] / ro;
                    const double vz = U[4] / ro;
                    const double vv = vx * vx + vy * vy + vz * vz;
                    const double bb = U[5] * U[5] + U[6] * U[6] + U[7] * U[7];
                    const double pr = (gm - 1.0) * (U[1] - 0.5 * ro * vv - pi8i * bb);
                    // the next step's cfl_m term (b2 and v2 are bb and vv,
                    // the same sums)
                    if (cxy && s >= mg + 1 && s <= kx - mg) {
                        const double dg = fminF(dxyg, image_min(a.dz, s, kx, mg));
                        const double dtq = cfl3_term(a.safety, gm, ro, pr, vv, bb, dg);
                        if (dtq < dtmin_local) dtmin_local = dtq;   // NaN never wins
                    }
                    a.out[0][pk] = ro;
                    a.out[1][pk] = pr;
                    a.out[2][pk] = vx;
                    a.out[3][pk] = vy;
                    a.out[4][pk] = vz;
                    a.out[5][pk] = U[5];
                    a.out[6][pk] = U[6];
                    a.out[7][pk] = U[7];
                }
                const double qzb = qav * zq * fmaxF(zero, __builtin_fabs(vzs - vzp) - kR4_1em4);
                const double vxc = ldv(L, t, par, L_VX), vyc = ldv(L, t, par, L_VY);
                const double qxc = qav * qx0 * fmaxF(zero, __builtin_fabs(ldv(L, t + 1, par, L_VX) - vxc) - kR4_1em4);
                const double qxw = qav * qx1 * fmaxF(zero, __builtin_fabs(vxc - ldv(L, t - 1, par, L_VX)) - kR4_1em4);
                const double qyc = qav * qy0 * fmaxF(zero, __builtin_fabs(ldv(L, t + TB, par, L_VY) - vyc) - kR4_1em4);
                const double qys = qav * qy1 * fmaxF(zero, __builtin_fabs(vyc - ldv(L, t - TB, par, L_VY)) - kR4_1em4);
                if (s < k1) {
#pragma unroll
                    for (int q = 0; q < 8; q++) {
                        double du = 0.0;
                        du = du - 0.5 * dt * (0.5 * dxi * (ldf(L, t + 1, par, q, 0) - ldf(L, t - 1, par, q, 0)))
                                - 0.5 * dt * (0.5 * dyi * (ldf(L, t + TB, par, q, 1) - ldf(L, t - TB, par, q, 1)));
                        dxy[q] = du;
                        const double ucq = ldu(L, t, par, q);
                        avx[q] = dt * (dxi * (qxc * dxim * (ldu(L, t + 1, par, q) - ucq)
                                              - qxw * dxim1 * (ucq - ldu(L, t - 1, par, q))));
                        avy[q] = dt * (dyi * (qyc * dyim * (ldu(L, t + TB, par, q) - ucq)
                                              - qys * dyim1 * (ucq - ldu(L, t - TB, par, q))));
                        bzl[q] = qzb * zdim * (ucq - ldu(L, t, lp, q));
                    }
                }
#pragma unroll
                for (int q = 0; q < 8; q++) fz2[q] = ldf(L, t, lp, q, 2);   // Fz of plane s-1
            }
            vzp = vzs;
        }
    }
    if (do_cfl) block_min_to_global(dtmin_local, red, &a.clk->dtkey[(a.step + 1) % 3]);
}

}  // namespace

hipError_t launch_mlw3d_march(const Mlw3dArgs &a, hipStream_t s) {
    if (a.ix < 3 || a.jx < 3 || a.kx < 3) return hipSuccess;
    static int ncu = 0;
    if (!ncu) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
            ncu = 256;
    }
    const int tiles_x = (a.ix - 2 + TO - 1) / TO, tiles_y = (a.jx - 2 + TO - 1) / TO;
    const int ntiles = tiles_x * tiles_y, nz = a.kx - 2;
    // z-segments: one block per CU (the LDS of a block), so a launch runs in
    // rounds of ncu blocks; take the number of segments with the fewest
    // plane-steps per CU, counting the two planes each segment marches
    // before its first update
    int nseg = 1;
    long best = -1;
    for (int n = 1; n <= 16 && n <= nz; n++) {
        const int len = (nz + n - 1) / n;
        const int used = (nz + len - 1) / len;
        const long rounds = ((long)ntiles * used + ncu - 1) / ncu;
        const long cost = rounds * (len + 2);
        if (best < 0 || cost < best) { best = cost; nseg = used; }
    }
    const int seglen = (nz + nseg - 1) / nseg;
    const long items = (long)ntiles * nseg;
    const unsigned grid = (unsigned)((items + kXcds - 1) / kXcds * kXcds);
    hipLaunchKernelGGL(mlw3d_march, dim3(grid), dim3(NT), 0, s, a, tiles_x, ntiles, nseg, seglen);
    return hipGetLastError();
}

}  // namespace cans
