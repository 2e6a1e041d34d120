// cans_mlw2d.hip -- fused 2D modified Lax-Wendroff MHD step (gfx950).
//
// One launch does the whole of cans2d/hdmlw/mlw_m.f:2-324 (mlwhalf, mlwsrch,
// mlwfull, mlwsrcf, mlwartv, update, primitive recovery), and optionally the
// periodic ghost fill of cans2d/bc/bdper{x,y}.f and the CFL reduction of
// cans2d/common/cfl_m.f for the next step.  State is structure-of-arrays
// FP64, column-major (ix,jx) as the reference holds it.
//
// Mapping: one wavefront owns a strip of 64 columns (62 output columns and
// one halo column each side) and marches along y over H output rows.  The
// stencil rows live in registers; x neighbours come from the adjacent lanes
// through DPP wave shifts (v_mov_b32_dpp wave_shl:1 / wave_shr:1), so there is
// no LDS traffic and no barrier, and each field is read once per strip and
// row with coalesced 512-byte wave loads.  The carried row state is kept to
// what the next row cannot re-derive by a lane shift (the east cell is
// shifted again rather than carried), so that a wave fits in 256 VGPRs and
// two waves share each SIMD: the FP64 chains of one hide the latency of the
// other's.
//
// Per output row j a lane forms the cell state of row j+1 (mlw_m.f:101-176),
// the corner half step (i,j) between rows j and j+1 (mlwhalf.f:45-52,
// mlw_m.f:189-262), and the update of cell (i,j).  Differences that the
// reference forms for two neighbouring cells are the same IEEE operation on
// the same operands -- e.g. (u(i+1)-u(i)) is uE-uc of cell i and uc-uW of
// cell i+1 (mlwartv.f:33-41), the corner-flux difference of corner row j is
// the north difference of row j and the south difference of row j+1
// (mlwfull.f:31-41) -- so each is computed once and handed to the neighbour
// by a lane shift or carried to the next row.  Every value is formed by the
// reference's expression in the reference's order, so the results are
// bit-identical to it and independent of the strip decomposition.
#include "cans_common.hpp"
#include "cans_clock.hpp"
#include "cans_kernels.hpp"
#include "cans_wave.hpp"

namespace cans {
namespace {

constexpr int WAVE = 64;
constexpr int COLS = WAVE - 2;               // output columns per wave
constexpr int WPB = 4;                       // waves per block
constexpr int kOccBatch = 2;                 // waves per SIMD of the batched kernel (the strip's registers)
constexpr double pi8i = 1.0 / kPi / 8.0;     // mlw_m.f:27  pi8i=1./pi/8.
constexpr double pi4i = 1.0 / kPi / 4.0;     // mlw_m.f:28  pi4i=1./pi/4.

// x neighbours: shl1/shr1 of cans_wave.hpp (lane l receives lane l+1 / l-1
// through a DPP wave shift; the edge lanes receive 0, and their results are
// never used by an output lane).

// u: ro ee rx ry bx by (the six conserved variables, mlw_m.f:101-112)
enum { U_RO, U_EE, U_RX, U_RY, U_BX, U_BY, NU };

struct Prim {
    double ro, pr, vx, vy, bx, by;
};

// Conserved variables and cell-centred fluxes of one cell (mlw_m.f:101-176).
// x fluxes of the six equations: rx, eex, rxx, ryx, 0, -ez
// y fluxes:                      ry, eey, rxy, ryy, ez, 0
struct Cell {
    double u[NU];
    double eex, eey, rxx, rxy, ryx, ryy, ez;
    double vx, vy;
};

// The part of the east neighbour a cell needs (no vy).
struct CellE {
    double u[NU];
    double eex, eey, rxx, rxy, ryx, ryy, ez;
    double vx;
};

struct Corner {
    double rox, roy, eex, eey, rxx, rxy, ryx, ryy, ez;
};

// Row state carried from output row j-1 to output row j.
struct RowState {
    Cell c;          // cell (i, j)
    double sy[5];    // y fluxes ry eey rxy ryy ez of cell (i, j-1)
    double dS[NU];   // u(i,j) - u(i,j-1)               (= uc - uS of row j)
    double qys;      // AV coefficient of face j-1/2
    double ky[5];    // y fluxes roy eey rxy ryy ez of corner (i, j-1)
    double dXS[5];   // corner x-flux differences of corner row j-1
    double PS;       // ux0*(-ezh(i-1,j-1)) + ux1*(-ezh(i,j-1))
    Prim nxt;        // primitives of cell (i, j+1), loaded
    double az;       // az of cell (i, j), loaded
    double pro, pee, prx, pry, pbx, pby, paz;   // updated cell (i, j)
};

// Per-row (y) constants, read through non-aliased pointers.
struct YMet {
    const double *dim, *di, *u0, *u1, *dym, *dy;
};

// Per-column constants of a strip.
struct ColCtx {
    int i;
    bool out;                     // output column
    unsigned boff;                // byte offset of column min(i,ix) within a row
    double dxi, dxim, dxim1;      // 1/dx(i), 1/dxm(i), 1/dxm(i-1)
    double ux0, ux1;              // 0.5*dxm(i)/dx(i), 0.5*dxm(i-1)/dx(i)
    double qxs;                   // qav*dxm(i)
    double dxc;                   // dx(i)
    double dxg;                   // min of dx(i) and dx of i's periodic x images
};

// Row j (clamped to the grid, so every lane reads a real cell) of a field:
// a wave-uniform base address plus the lane's 32-bit byte offset.
CANS_DEV const double *row_of(const double *f, int ix, int jx, int j) {
    const int jc = j < 1 ? 1 : (j > jx ? jx : j);
    return f + (size_t)(jc - 1) * (size_t)ix;
}
CANS_DEV double ld(const double *row, unsigned boff) {
    return *reinterpret_cast<const double *>(reinterpret_cast<const char *>(row) + boff);
}
CANS_DEV void st(double *row, unsigned boff, double v) {
    *reinterpret_cast<double *>(reinterpret_cast<char *>(row) + boff) = v;
}

CANS_DEV Prim load_prim(const Mlw2dArgs &a, const ColCtx &x, int j) {
    const size_t r = row_of(a.ro, a.ix, a.jx, j) - a.ro;
    // Loads as scalar row base + 32-bit lane offset; the empty asm keeps the
    // compiler from hoisting six 64-bit per-lane field pointers into VGPRs.
    unsigned bo = x.boff;
    asm volatile("" : "+v"(bo));
    Prim p;
    p.ro = ld(a.ro + r, bo); p.pr = ld(a.pr + r, bo); p.vx = ld(a.vx + r, bo);
    p.vy = ld(a.vy + r, bo); p.bx = ld(a.bx + r, bo); p.by = ld(a.by + r, bo);
    return p;
}

CANS_DEV double load_az(const Mlw2dArgs &a, const ColCtx &x, int j) {
    unsigned bo = x.boff;
    asm volatile("" : "+v"(bo));
    return ld(row_of(a.az, a.ix, a.jx, j), bo);
}

// mlw_m.f:101-176 for one cell
CANS_DEV Cell make_cell(const Prim &p, double gm) {
    Cell c;
    const double ro = p.ro, pr = p.pr, vx = p.vx, vy = p.vy, bx = p.bx, by = p.by;
    const double vv = vx * vx + vy * vy;
    const double bb = bx * bx + by * by;
    const double ez = -vx * by + vy * bx;
    const double ep = pr * gm / (gm - 1.0) + 0.5 * ro * vv;
    c.u[U_RO] = ro;
    c.u[U_EE] = pr / (gm - 1.0) + 0.5 * ro * vv + pi8i * bb;
    c.u[U_RX] = ro * vx;
    c.u[U_RY] = ro * vy;
    c.u[U_BX] = bx;
    c.u[U_BY] = by;
    c.eex = ep * vx - by * ez * pi4i;
    c.eey = ep * vy + bx * ez * pi4i;
    c.rxx = ro * (vx * vx) + pr + pi8i * (by * by - bx * bx);
    c.rxy = ro * vx * vy - pi4i * bx * by;
    c.ryx = ro * vy * vx - pi4i * by * bx;
    c.ryy = ro * (vy * vy) + pr + pi8i * (bx * bx - by * by);
    c.ez = ez;
    c.vx = vx;
    c.vy = vy;
    return c;
}

CANS_DEV CellE shl_cell(const Cell &c) {
    CellE r;
#pragma unroll
    for (int q = 0; q < NU; q++) r.u[q] = shl1(c.u[q]);
    r.eex = shl1(c.eex); r.eey = shl1(c.eey); r.rxx = shl1(c.rxx); r.rxy = shl1(c.rxy);
    r.ryx = shl1(c.ryx); r.ryy = shl1(c.ryy); r.ez = shl1(c.ez); r.vx = shl1(c.vx);
    return r;
}

// one conserved component of the corner half step (mlwhalf.f:47-50):
// A=(i,j) B=(i+1,j) C=(i,j+1) D=(i+1,j+1)
CANS_DEV double half1(double uA, double uB, double uC, double uD, double fA, double fB, double fC, double fD,
                      double gA, double gB, double gC, double gD, double dxim, double dyim, double dt) {
    const double uh = 0.25 * (uB + uA + uD + uC);
    const double dfdx = dxim * (fB - fA + fD - fC) / 2.0;
    const double dfdy = dyim * (gC - gA + gD - gB) / 2.0;
    return uh - dt * dfdx - dt * dfdy;
}

// bx has no x flux and by no y flux.  The reference still forms the zero
// term, dt*(dm*(+0)/2) = +0 for finite dt >= 0 and positive metric, and
// subtracting +0 leaves every value (-0 included) unchanged: these variants
// omit it and give the same bits.
CANS_DEV double half1_y(double uA, double uB, double uC, double uD, double gA, double gB, double gC, double gD,
                        double dyim, double dt) {
    const double uh = 0.25 * (uB + uA + uD + uC);
    const double dfdy = dyim * (gC - gA + gD - gB) / 2.0;
    return uh - dt * dfdy;
}
CANS_DEV double half1_x(double uA, double uB, double uC, double uD, double fA, double fB, double fC, double fD,
                        double dxim, double dt) {
    const double uh = 0.25 * (uB + uA + uD + uC);
    const double dfdx = dxim * (fB - fA + fD - fC) / 2.0;
    return uh - dt * dfdx;
}

// corner between columns i,i+1 and rows j,j+1 (mlwhalf.f:45-52, mlw_m.f:189-262)
template <class CA, class CB>
CANS_DEV Corner make_corner(const CA &A, const CB &B, const CA &C, const CB &D, double dxim, double dyim, double dt,
                            double gm) {
    const double roh = half1(A.u[U_RO], B.u[U_RO], C.u[U_RO], D.u[U_RO], A.u[U_RX], B.u[U_RX], C.u[U_RX],
                             D.u[U_RX], A.u[U_RY], B.u[U_RY], C.u[U_RY], D.u[U_RY], dxim, dyim, dt);
    const double eeh = half1(A.u[U_EE], B.u[U_EE], C.u[U_EE], D.u[U_EE], A.eex, B.eex, C.eex, D.eex, A.eey, B.eey,
                             C.eey, D.eey, dxim, dyim, dt);
    const double rxh = half1(A.u[U_RX], B.u[U_RX], C.u[U_RX], D.u[U_RX], A.rxx, B.rxx, C.rxx, D.rxx, A.rxy, B.rxy,
                             C.rxy, D.rxy, dxim, dyim, dt);
    const double ryh = half1(A.u[U_RY], B.u[U_RY], C.u[U_RY], D.u[U_RY], A.ryx, B.ryx, C.ryx, D.ryx, A.ryy, B.ryy,
                             C.ryy, D.ryy, dxim, dyim, dt);
    const double bxh = half1_y(A.u[U_BX], B.u[U_BX], C.u[U_BX], D.u[U_BX], A.ez, B.ez, C.ez, D.ez, dyim, dt);
    const double byh = half1_x(A.u[U_BY], B.u[U_BY], C.u[U_BY], D.u[U_BY], -A.ez, -B.ez, -C.ez, -D.ez, dxim, dt);
    const double vxh = rxh / roh;
    const double vyh = ryh / roh;
    const double vv = vxh * vxh + vyh * vyh;
    const double bb = bxh * bxh + byh * byh;
    const double prh = (gm - 1.0) * (eeh - 0.5 * roh * vv - pi8i * bb);
    const double ezh = -vxh * byh + vyh * bxh;
    const double ep = prh * gm / (gm - 1.0) + 0.5 * roh * vv;
    Corner k;
    k.rox = roh * vxh;
    k.roy = roh * vyh;
    k.eex = ep * vxh - byh * ezh * pi4i;
    k.eey = ep * vyh + bxh * ezh * pi4i;
    k.rxx = roh * (vxh * vxh) + prh + pi8i * (byh * byh - bxh * bxh);
    k.rxy = roh * vxh * vyh - pi4i * bxh * byh;
    k.ryx = roh * vyh * vxh - pi4i * byh * bxh;
    k.ryy = roh * (vyh * vyh) + prh + pi8i * (bxh * bxh - byh * byh);
    k.ez = ezh;
    return k;
}

CANS_DEV double av_coef(double qs, double vhi, double vlo) {   // mlw_m.f:276-288
    const double zero = 0.0;
    return qs * fmaxF(zero, __builtin_fabs(vhi - vlo) - kR4_1em4);
}

// du of one equation: centre term of mlwhalf (mlwhalf.f:36-41), mlwfull
// (mlwfull.f:31-41) and mlwartv (mlwartv.f:33-41), in the reference's order.
//   dfx = fE-fW   dgy = gN-gS   hS/hN = corner x differences south/north
//   kW/kE = corner y differences west/east   dE = uE-uc  dW = uc-uW
//   dN = uN-uc   dS = uc-uS
struct RowCoef {
    double hdt, dt;               // 0.5*dt, dt
    double hdxi, dxi, hdyi, dyi;  // 0.5/dx(i), 1/dx(i), 0.5/dy(j), 1/dy(j)
    double ux0, ux1, uy0, uy1;
    double qxe, qxw, qyn, qys;    // qxc*dxim, qxw*dxim1, qyc*dyim, qys*dyim1
};

CANS_DEV double du1(const RowCoef &r, double dfx, double dgy, double hS, double hN, double kW, double kE, double dE,
                    double dW, double dN, double dS) {
    double du = 0.0;
    du = du - r.hdt * (r.hdxi * dfx) - r.hdt * (r.hdyi * dgy);
    const double dfdx = r.dxi * (r.uy0 * hS + r.uy1 * hN);
    const double dfdy = r.dyi * (r.ux0 * kW + r.ux1 * kE);
    du = du - r.hdt * dfdx - r.hdt * dfdy;
    du = du + r.dt * (r.dxi * (r.qxe * dE - r.qxw * dW)) + r.dt * (r.dyi * (r.qyn * dN - r.qys * dS));
    return du;
}

// bx (no x flux) and by (no y flux): the zero-flux terms of du1 are +0 for
// finite dt >= 0 and positive metric and are omitted (same bits).
CANS_DEV double du1_y(const RowCoef &r, double dgy, double kW, double kE, double dE, double dW, double dN,
                      double dS) {
    double du = 0.0;
    du = du - r.hdt * (r.hdyi * dgy);
    const double dfdy = r.dyi * (r.ux0 * kW + r.ux1 * kE);
    du = du - r.hdt * dfdy;
    du = du + r.dt * (r.dxi * (r.qxe * dE - r.qxw * dW)) + r.dt * (r.dyi * (r.qyn * dN - r.qys * dS));
    return du;
}
CANS_DEV double du1_x(const RowCoef &r, double dfx, double hS, double hN, double dE, double dW, double dN,
                      double dS) {
    double du = 0.0;
    du = du - r.hdt * (r.hdxi * dfx);
    const double dfdx = r.dxi * (r.uy0 * hS + r.uy1 * hN);
    du = du - r.hdt * dfdx;
    du = du + r.dt * (r.dxi * (r.qxe * dE - r.qxw * dW)) + r.dt * (r.dyi * (r.qyn * dN - r.qys * dS));
    return du;
}

// One equation with both an x and a y flux (ro ee rx ry).  f*: cell x flux,
// g*: cell y flux, h*: corner x flux, k*: corner y flux.  The differences
// handed on: hN (south difference of the next row), kE (shifted to the east
// neighbour as its kW), dN (the next row's dS).
CANS_DEV double eq_xy(const RowCoef &r, double fE, double fC, double gN, double gS, double hC, double hS, double kC,
                      double kS, double uE, double uc, double uN, double dS, double &hN, double &dN) {
    const double dfx = fE - shr1(fC);
    const double dgy = gN - gS;
    hN = hC - shr1(hC);
    const double kE = kC - kS;
    const double kW = shr1(kE);
    const double dE = uE - uc;
    const double dW = shr1(dE);
    dN = uN - uc;
    return du1(r, dfx, dgy, hS, hN, kW, kE, dE, dW, dN, dS);
}

// Primitive recovery (mlw_m.f:299-321), the CFL term of the new state
// (cans2d/common/cfl_m.f:46-55) and the stores of the updated cell (i, j),
// with its periodic ghost images.  When the kernel writes the ghost images
// (per_margin = m > 0) only interior cells, i in [m+1, ix-m] and j in
// [m+1, jx-m], store and enter the CFL: the reference's bdperx/bdpery
// overwrite what mlw_m computed in the ghost cells 2..m and ix-m+1..ix-1
// with interior images.
CANS_DEV void finish_row(const Mlw2dArgs &a, const YMet &y, const ColCtx &x, int j, const RowState &s,
                         bool do_cfl, double &dtmin_local) {
    const int mgj = a.per_margin;
    const bool out = x.out && (mgj == 0 || (j >= mgj + 1 && j <= a.jx - mgj));   // uniform in j
    const double gm = a.gm;
    const double ro = s.pro, ee = s.pee, rx = s.prx, ry = s.pry, bx = s.pbx, by = s.pby, az = s.paz;
    const double vx = rx / ro;
    const double vy = ry / ro;
    const double vv = vx * vx + vy * vy;
    const double bb = bx * bx + by * by;
    const double pr = (gm - 1.0) * (ee - 0.5 * ro * vv - pi8i * bb);
    double sq = 0.0;
    if (do_cfl) {
        const double b2 = bx * bx + by * by;
        const double ca2 = b2 * pi4i / ro;
        const double cs2 = gm * pr / ro;
        const double v2 = vx * vx + vy * vy;
        sq = __builtin_sqrt(v2 + cs2 + ca2);
        // The reference's cfl_m also visits the periodic ghost images of
        // this cell, which hold the same state: their terms are
        // safety*(min(dx(gi), dy(gj))/sq) over the images (gi, gj).  For
        // sq >= 0, m -> safety*(m/sq) is monotone non-decreasing in IEEE
        // arithmetic (a NaN sq gives NaN terms that never win either way), so
        // the smallest of those terms is safety*(min(min dx, min dy)/sq) bit
        // for bit: one division per cell and no per-image metric loads.
        const int mg = a.per_margin, jx = a.jx;
        double dyg = y.dy[j - 1];
        if (mg > 0 && j >= jx - 2 * mg + 1) dyg = fminF(dyg, y.dy[j - (jx - 2 * mg) - 1]);
        if (mg > 0 && j <= 2 * mg) dyg = fminF(dyg, y.dy[j + (jx - 2 * mg) - 1]);
        const double dtq = a.safety * (fminF(x.dxg, dyg) / sq);
        dtmin_local = (out && dtq < dtmin_local) ? dtq : dtmin_local;
    }
    if (out) {
        const int ix = a.ix, jx = a.jx;
        // Stores as scalar row address + 32-bit lane offset.  The empty asm
        // makes the offset opaque so that the compiler does not hoist seven
        // 64-bit per-lane field pointers out of the loop into registers.
        const size_t rw = (size_t)(__builtin_amdgcn_readfirstlane(j) - 1) * (size_t)ix;
        unsigned bo = x.boff;
        asm volatile("" : "+v"(bo));
        st(a.ro_o + rw, bo, ro); st(a.pr_o + rw, bo, pr); st(a.vx_o + rw, bo, vx);
        st(a.vy_o + rw, bo, vy); st(a.bx_o + rw, bo, bx); st(a.by_o + rw, bo, by);
        st(a.az_o + rw, bo, az);
        // In a periodic box also the ghost images: after bdperx then
        // bdpery, ghost (gi,gj) holds interior (src_x(gi), src_y(gj)).
        const int mg = a.per_margin;   // the host checks ix-2m >= m and jx-2m >= m
        const int i = x.i;
        if (mg > 0 && (i >= ix - 2 * mg + 1 || i <= 2 * mg || j >= jx - 2 * mg + 1 || j <= 2 * mg)) {
            int xi[3] = {i, 0, 0}, yj[3] = {j, 0, 0};
            int nxi = 1, nyj = 1;
            if (i >= ix - 2 * mg + 1) xi[nxi++] = i - (ix - 2 * mg);   // low ghost  <- high interior
            if (i <= 2 * mg) xi[nxi++] = i + (ix - 2 * mg);            // high ghost <- low interior
            if (j >= jx - 2 * mg + 1) yj[nyj++] = j - (jx - 2 * mg);
            if (j <= 2 * mg) yj[nyj++] = j + (jx - 2 * mg);
            for (int b = 0; b < nyj; b++) {
                for (int c = 0; c < nxi; c++) {
                    if (b == 0 && c == 0) continue;
                    const size_t q = (size_t)(xi[c] - 1) + (size_t)ix * (size_t)(yj[b] - 1);
                    a.ro_o[q] = ro; a.pr_o[q] = pr; a.vx_o[q] = vx; a.vy_o[q] = vy;
                    a.bx_o[q] = bx; a.by_o[q] = by; a.az_o[q] = az;
                }
            }
        }
    }
}

// One output row j: s holds the state entering row j, o receives the state
// entering row j+1.  Everything but the stores runs on all 64 lanes (the
// lane shifts need every lane's operands).
CANS_DEV void row_step(const Mlw2dArgs &a, const YMet &y, const ColCtx &x, int j, double dt, bool do_cfl,
                       const RowState &s, RowState &o, double &dtmin_local) {
    j = __builtin_amdgcn_readfirstlane(j);   // wave-uniform: SGPR row bases and scalar loads
    const double gm = a.gm;
    // prefetch the row after next and the az of the next row
    o.nxt = load_prim(a, x, j + 2);
    o.az = load_az(a, x, j + 1);

    const double dyim = y.dim[j - 1];
    const CellE E = shl_cell(s.c);                 // cell (i+1, j)
    const Cell R2 = make_cell(s.nxt, gm);          // cell (i, j+1)
    const CellE R2p = shl_cell(R2);                // cell (i+1, j+1)
    const Corner C = make_corner(s.c, E, R2, R2p, x.dxim, dyim, dt, gm);

    const double qxc = av_coef(x.qxs, E.vx, s.c.vx);
    const double qyc = av_coef(a.qav * y.dym[j - 1], R2.vy, s.c.vy);
    RowCoef r;
    r.hdt = 0.5 * dt;
    r.dt = dt;
    r.dxi = x.dxi;
    r.hdxi = 0.5 * x.dxi;
    r.dyi = y.di[j - 1];
    r.hdyi = 0.5 * r.dyi;
    r.ux0 = x.ux0; r.ux1 = x.ux1;
    r.uy0 = y.u0[j - 1]; r.uy1 = y.u1[j - 1];
    r.qxe = qxc * x.dxim;
    r.qxw = shr1(qxc) * x.dxim1;
    r.qyn = qyc * dyim;
    r.qys = s.qys * y.dim[j - 2];

    double du[NU];
    du[U_RO] = eq_xy(r, E.u[U_RX], s.c.u[U_RX], R2.u[U_RY], s.sy[0], C.rox, s.dXS[0], C.roy, s.ky[0],
                     E.u[U_RO], s.c.u[U_RO], R2.u[U_RO], s.dS[U_RO], o.dXS[0], o.dS[U_RO]);
    du[U_EE] = eq_xy(r, E.eex, s.c.eex, R2.eey, s.sy[1], C.eex, s.dXS[1], C.eey, s.ky[1], E.u[U_EE],
                     s.c.u[U_EE], R2.u[U_EE], s.dS[U_EE], o.dXS[1], o.dS[U_EE]);
    du[U_RX] = eq_xy(r, E.rxx, s.c.rxx, R2.rxy, s.sy[2], C.rxx, s.dXS[2], C.rxy, s.ky[2], E.u[U_RX],
                     s.c.u[U_RX], R2.u[U_RX], s.dS[U_RX], o.dXS[2], o.dS[U_RX]);
    du[U_RY] = eq_xy(r, E.ryx, s.c.ryx, R2.ryy, s.sy[3], C.ryx, s.dXS[3], C.ryy, s.ky[3], E.u[U_RY],
                     s.c.u[U_RY], R2.u[U_RY], s.dS[U_RY], o.dXS[3], o.dS[U_RY]);
    const double ezW = shr1(C.ez);                 // ezh(i-1, j)
    {   // bx: x flux 0, y flux ez
        const double kE = C.ez - s.ky[4];
        const double dE = E.u[U_BX] - s.c.u[U_BX];
        const double dN = R2.u[U_BX] - s.c.u[U_BX];
        du[U_BX] = du1_y(r, R2.ez - s.sy[4], shr1(kE), kE, dE, shr1(dE), dN, s.dS[U_BX]);
        o.dS[U_BX] = dN;
    }
    {   // by: x flux -ez, y flux 0.  (-a)-(-b) is b-a bit for bit (signed
        // zeros included), so the x differences are taken as west - east.
        const double hN = ezW - C.ez;
        const double dE = E.u[U_BY] - s.c.u[U_BY];
        const double dN = R2.u[U_BY] - s.c.u[U_BY];
        du[U_BY] = du1_x(r, shr1(s.c.ez) - E.ez, s.dXS[4], hN, dE, shr1(dE), dN, s.dS[U_BY]);
        o.dXS[4] = hN;
        o.dS[U_BY] = dN;
    }
    // Az: mlwsrch.f:32-36 then mlwsrcf.f:33-40, source s = -ez
    const double PN = x.ux0 * (-ezW) + x.ux1 * (-C.ez);
    double daz = 0.0;
    daz = daz + r.hdt * (-s.c.ez);
    const double sh = r.uy0 * s.PS + r.uy1 * PN;
    daz = daz + r.hdt * sh;
    // update (mlw_m.f:299-304)
    o.pro = s.c.u[U_RO] + du[U_RO];
    o.pee = s.c.u[U_EE] + du[U_EE];
    o.prx = s.c.u[U_RX] + du[U_RX];
    o.pry = s.c.u[U_RY] + du[U_RY];
    o.pbx = s.c.u[U_BX] + du[U_BX];
    o.pby = s.c.u[U_BY] + du[U_BY];
    o.paz = s.az + daz;

    // state entering row j+1
    o.c = R2;
    o.sy[0] = s.c.u[U_RY]; o.sy[1] = s.c.eey; o.sy[2] = s.c.rxy; o.sy[3] = s.c.ryy; o.sy[4] = s.c.ez;
    o.qys = qyc;
    o.ky[0] = C.roy; o.ky[1] = C.eey; o.ky[2] = C.rxy; o.ky[3] = C.ryy; o.ky[4] = C.ez;
    o.PS = PN;

    finish_row(a, y, x, j, o, do_cfl, dtmin_local);
}

// One wave's strip: columns 1+strip*COLS .. (lane 0 the west halo), output
// rows jb = 2+rowb*H .. min(jb+H-1, jx-1), marched in y (prologue: the cells
// of rows jb-1 and jb and the corner row jb-1).
CANS_DEV void march_strip(const Mlw2dArgs &a, const YMet &y, int H, int strip, int rowb, double dt, bool do_cfl,
                          double &dtmin_local) {
    const int ix = a.ix, jx = a.jx;
    const int lane = threadIdx.x & (WAVE - 1);
    const int jb = 2 + rowb * H;
    const int je = min(jb + H - 1, jx - 1);
    // Issue priority by progress: two waves share a SIMD, and with equal
    // priorities the older one issues first, finishes early and leaves the
    // younger one to end its rows alone, with its FP64 latencies exposed.  A
    // wave starts at priority 3 and steps down by one per quarter of its
    // rows, so the one further behind issues first and the pair ends
    // together (1024^2 step 38.0 -> 36.1 us, 4096^2 497 -> 488 us).
    __builtin_amdgcn_s_setprio(3);
    ColCtx x;
    x.i = 1 + strip * COLS + lane;         // lane 0 is the west halo column
    x.out = lane >= 1 && lane <= COLS && x.i <= ix - 1 &&
            (a.per_margin == 0 || (x.i >= a.per_margin + 1 && x.i <= ix - a.per_margin));
    const int ic = min(x.i, ix);           // lanes past the grid read column ix
    x.boff = (unsigned)(ic - 1) * 8u;
    x.dxi = a.mx.di[ic - 1];
    x.dxim = a.mx.dim[ic - 1];
    x.dxim1 = a.mx.dim[max(ic - 2, 0)];
    x.ux0 = a.mx.u0[ic - 1];
    x.ux1 = a.mx.u1[ic - 1];
    x.qxs = a.qav * a.dxm[ic - 1];
    x.dxc = a.dx[ic - 1];
    x.dxg = x.dxc;
    {
        const int mg = a.per_margin;
        if (mg > 0 && ic >= ix - 2 * mg + 1) x.dxg = fminF(x.dxg, a.dx[ic - (ix - 2 * mg) - 1]);
        if (mg > 0 && ic <= 2 * mg) x.dxg = fminF(x.dxg, a.dx[ic + (ix - 2 * mg) - 1]);
    }

    // prologue: cells of rows jb-1 and jb, corner row jb-1
    const Cell R0 = make_cell(load_prim(a, x, jb - 1), a.gm);
    RowState s0, s1;
    s0.c = make_cell(load_prim(a, x, jb), a.gm);
    const CellE s0e = shl_cell(s0.c);
    s0.nxt = load_prim(a, x, jb + 1);
    s0.az = load_az(a, x, jb);
    const CellE R0p = shl_cell(R0);
    const Corner Cm = make_corner(R0, R0p, s0.c, s0e, x.dxim, a.my.dim[jb - 2], dt, a.gm);
    const double ezW = shr1(Cm.ez);
    s0.dXS[0] = Cm.rox - shr1(Cm.rox);
    s0.dXS[1] = Cm.eex - shr1(Cm.eex);
    s0.dXS[2] = Cm.rxx - shr1(Cm.rxx);
    s0.dXS[3] = Cm.ryx - shr1(Cm.ryx);
    s0.dXS[4] = ezW - Cm.ez;
    s0.PS = x.ux0 * (-ezW) + x.ux1 * (-Cm.ez);
    s0.ky[0] = Cm.roy; s0.ky[1] = Cm.eey; s0.ky[2] = Cm.rxy; s0.ky[3] = Cm.ryy; s0.ky[4] = Cm.ez;
    s0.sy[0] = R0.u[U_RY]; s0.sy[1] = R0.eey; s0.sy[2] = R0.rxy; s0.sy[3] = R0.ryy; s0.sy[4] = R0.ez;
#pragma unroll
    for (int q = 0; q < NU; q++) s0.dS[q] = s0.c.u[q] - R0.u[q];
    s0.qys = av_coef(a.qav * a.dym[jb - 2], s0.c.vy, R0.vy);

    // two rows per trip so that the carried state ping-pongs between
    // s0 and s1 instead of being copied
    int j = jb;
    const int nr = je - jb + 1;
    for (; j + 1 <= je; j += 2) {
        const int q4 = (4 * (j - jb)) / nr;   // quarters of the rows done
        if (q4 == 1) __builtin_amdgcn_s_setprio(2);
        else if (q4 == 2) __builtin_amdgcn_s_setprio(1);
        else if (q4 >= 3) __builtin_amdgcn_s_setprio(0);
        row_step(a, y, x, j, dt, do_cfl, s0, s1, dtmin_local);
        row_step(a, y, x, j + 1, dt, do_cfl, s1, s0, dtmin_local);
    }
    if (j <= je) row_step(a, y, x, j, dt, do_cfl, s0, s1, dtmin_local);
}

// The y metric arrays come as separate non-aliased arguments (the same
// pointers as a.my, a.dym, a.dy) so that their per-row uniform loads go
// through the scalar cache.
template <int OCC>
__global__ __launch_bounds__(WAVE * WPB) __attribute__((amdgpu_waves_per_eu(OCC))) void mlw2d_kernel(
    Mlw2dArgs a, int H, const double *__restrict__ ydim, const double *__restrict__ ydi,
    const double *__restrict__ yu0, const double *__restrict__ yu1, const double *__restrict__ ydym,
    const double *__restrict__ ydy) {
    const YMet y{ydim, ydi, yu0, yu1, ydym, ydy};
    __shared__ double red[WPB];
    const int ix = a.ix, jx = a.jx;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nstrip = (ix - 2 + COLS - 1) / COLS;
    const int nrowb = (jx - 2 + H - 1) / H;
    const int nblocks = (nstrip * nrowb + WPB - 1) / WPB;
    const int unit = xcd_remap(blockIdx.x, nblocks) * WPB + wv;

    double dt = a.dt;
    bool run = true;
    if (a.clk) {
        StepCtl c = clock_use(a.clk, a.step, a.peek != 0, blockIdx.x == 0 && threadIdx.x == 0);
        run = c.run;
        dt = c.dt;
    }
    if (!run) return;                          // uniform across the launch
    if (a.dt_log && blockIdx.x == 0 && threadIdx.x == 0) a.dt_log[a.step] = dt;
    const bool do_cfl = a.clk && a.safety > 0.0;
    double dtmin_local = kR4_1e20;

    if (unit < nstrip * nrowb) march_strip(a, y, H, unit % nstrip, unit / nstrip, dt, do_cfl, dtmin_local);
    if (do_cfl) block_min_to_global(dtmin_local, red, &a.clk->dtkey[(a.step + 1) % 3]);
}

// ---------------------------------------------------------------- batched runs
// Many independent periodic 2D runs, each running the cans2d main loop
// (cans2d/md_mhdshktb/hdmlw/main.f:150-206 with bdperx/bdpery):
//     ns=ns+1; cfl_m; stop on merr; t=t+dt; mlw_m; bdperx; bdpery;
//     repeat while ns < nstop and t < tend
// to its own end, in one persistent kernel.  A workgroup takes a run from a
// queue and advances it to the end before taking the next; inside a run its
// waves share the strips of each step (the register-marching strip of the
// single-run kernel, with the ghost images and the next step's cfl term
// fused), and a step ends with one workgroup barrier, the dt min-reduction.
// Writes of one step are read by other waves of the same workgroup in the
// next, after that barrier (one CU, one L1).

// rows per strip for a run of jx rows and nstrip strips on NW waves: the
// fewest rounds of units, each unit's prologue counted as 1.5 rows
CANS_DEV int batch2d_rows(int jx, int nstrip, int NW) {
    int best = 8;
    double cost = 1e300;
    for (int h = 4; h <= 48; h++) {
        const int units = nstrip * ((jx - 2 + h - 1) / h);
        const double c = (double)((units + NW - 1) / NW) * (h + 1.5);
        if (c < cost) { cost = c; best = h; }
    }
    return best;
}

template <int BT>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(kOccBatch))) void batch2d_kernel(Batch2dArgs b) {
    constexpr int NW = BT / 64;
    __shared__ double red[2][NW];
    __shared__ int run_s;
    const int wv = uni((int)(threadIdx.x >> 6));
    for (;;) {
        // queue pop by wave 0 under a uniform branch (lane 0 adds 1)
        if (wv == 0) {
            const unsigned got = atomicAdd(b.next, threadIdx.x == 0 ? 1u : 0u);
            run_s = uni((int)got);
        }
        __syncthreads();
        const int kq = uni(run_s);
        __syncthreads();
        if (kq >= b.nrun) break;
        const int k = b.order ? uni(b.order[kq]) : kq;
        const long long o = uni(b.off[k]);
        // step parity p reads set p and writes the other: runs[2k+p]
        const Mlw2dArgs &a0 = b.runs[2 * k];
        const int ix = uni(a0.ix), jx = uni(a0.jx);
        const size_t ncell = (size_t)ix * (size_t)jx;
        // cfl_m of the initial state over every cell (cans2d/common/cfl_m.f:44-62)
        double m = kR4_1e20;
        for (size_t p = threadIdx.x; p < ncell; p += BT) {
            const int i = (int)(p % (size_t)ix) + 1, j = (int)(p / (size_t)ix) + 1;
            const size_t q = (size_t)o + p;
            const double ro = b.f[0][q], pr = b.f[1][q], vx = b.f[2][q], vy = b.f[3][q], bx = b.f[4][q],
                         by = b.f[5][q];
            const double b2 = bx * bx + by * by;
            const double ca2 = b2 * pi4i / ro;
            const double cs2 = b.gm * pr / ro;
            const double v2 = vx * vx + vy * vy;
            const double dtq = b.safety * (fminF(a0.dx[i - 1], a0.dy[j - 1]) / __builtin_sqrt(v2 + cs2 + ca2));
            if (dtq < m) m = dtq;
        }
        int par = 0;
        double dt = block_min<BT>(m, red[par]);
        par ^= 1;
        int ns = 0, merr = 0, nexec = 0;
        double t = 0.0;
        const int nstrip = (ix - 2 + COLS - 1) / COLS;
        const int H = batch2d_rows(jx, nstrip, NW);
        const int nunit = nstrip * ((jx - 2 + H - 1) / H);
        for (;;) {
            ns = ns + 1;
            if (dt < b.dtmin) { merr = 9001; break; }     // cfl_m.f: merr=9001, main.f stops
            t = t + dt;
            const Mlw2dArgs &a = b.runs[2 * k + (nexec & 1)];
            const YMet y{a.my.dim, a.my.di, a.my.u0, a.my.u1, a.dym, a.dy};
            double dmin = kR4_1e20;
            for (int u = wv; u < nunit; u += NW) march_strip(a, y, H, u % nstrip, u / nstrip, dt, true, dmin);
            nexec++;
            dt = block_min<BT>(dmin, red[par]);           // dt of the next step; also the step's barrier
            par ^= 1;
            if (!(ns < b.nstop && t < b.tend)) break;
        }
        if (nexec & 1) {   // the last state is in the scratch set: copy back
            for (size_t p = threadIdx.x; p < ncell; p += BT)
#pragma unroll
                for (int q = 0; q < 7; q++) b.f[q][(size_t)o + p] = b.g[q][(size_t)o + p];
        }
        if (wv == 0) {   // uniform branch; the lanes store identical values
            b.t_out[k] = t;
            b.ns_out[k] = ns;
            b.merr_out[k] = merr;
        }
    }
}

// per-run metric of both axes (the arithmetic of the single-run metric
// kernel) and the queue counter; block r fills run r
__global__ void batch2d_prep(Batch2dArgs b) {
    const int r = blockIdx.x;
    if (r == 0 && threadIdx.x == 0) *b.next = 0u;
    if (r >= b.nrun) return;
    for (int axis = 0; axis < 2; axis++) {
        const long long o = axis ? b.offy[r] : b.offx[r];
        const int n = axis ? b.jx[r] : b.ix[r];
        const double *d = (axis ? b.dy : b.dx) + o, *dm = (axis ? b.dym : b.dxm) + o;
        double *const *m = axis ? b.my : b.mx;
        for (int i = 1 + (int)threadIdx.x; i <= n; i += blockDim.x) {
            m[0][o + i - 1] = 1.0 / d[i - 1];
            m[1][o + i - 1] = 1.0 / dm[i - 1];
            if (i >= 2 && i <= n - 1) {
                m[2][o + i - 1] = 0.5 * dm[i - 1] / d[i - 1];
                m[3][o + i - 1] = 0.5 * dm[i - 2] / d[i - 1];
            } else {
                m[2][o + i - 1] = 0.0;
                m[3][o + i - 1] = 0.0;
            }
        }
    }
}

// Build-time tuning: waves per SIMD the register budget is compiled for,
// and a fixed rows-per-wave (0: chosen per grid below).
#ifndef CANS_MLW2D_OCC
#define CANS_MLW2D_OCC 2
#endif
#ifndef CANS_MLW2D_ROWS
#define CANS_MLW2D_ROWS 0
#endif
constexpr int kOcc = CANS_MLW2D_OCC;
constexpr int kSlots = 256 * 4 * kOcc; // wave slots of the 256 CUs

// Rows per wave: each wave pays about 1.5 rows of prologue, and the launch
// runs in ceil(waves / slots) rounds, so pick H minimising rounds * (H + 1.5).
int pick_rows(int ix, int jx) {
    if (CANS_MLW2D_ROWS > 0) return CANS_MLW2D_ROWS;
    const int nstrip = (ix - 2 + COLS - 1) / COLS;
    int best = 16;
    double cost = 1e300;
    for (int h = 8; h <= 96; h++) {
        const long waves = (long)nstrip * ((jx - 2 + h - 1) / h);
        const long rounds = (waves + kSlots - 1) / kSlots;
        const double c = (double)rounds * (h + 1.5);
        if (c < cost) { cost = c; best = h; }
    }
    return best;
}

}  // namespace

hipError_t launch_batch2d(const Batch2dArgs &b, int block_threads, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(batch2d_prep, dim3(b.nrun > 0 ? b.nrun : 1), dim3(256), 0, s, b);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || b.nrun == 0) return e;
    switch (block_threads) {
    case 64: hipLaunchKernelGGL(batch2d_kernel<64>, dim3(nblocks), dim3(64), 0, s, b); break;
    case 128: hipLaunchKernelGGL(batch2d_kernel<128>, dim3(nblocks), dim3(128), 0, s, b); break;
    case 256: hipLaunchKernelGGL(batch2d_kernel<256>, dim3(nblocks), dim3(256), 0, s, b); break;
    case 512: hipLaunchKernelGGL(batch2d_kernel<512>, dim3(nblocks), dim3(512), 0, s, b); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_mlw2d(const Mlw2dArgs &a, hipStream_t s) {
    if (a.ix < 3 || a.jx < 3) return hipSuccess;
    const int h = pick_rows(a.ix, a.jx);
    const int nstrip = (a.ix - 2 + COLS - 1) / COLS;
    const int nrowb = (a.jx - 2 + h - 1) / h;
    const int nblocks = (nstrip * nrowb + WPB - 1) / WPB;
    hipLaunchKernelGGL(mlw2d_kernel<kOcc>, dim3(nblocks), dim3(WAVE * WPB), 0, s, a, h, a.my.dim, a.my.di, a.my.u0,
                       a.my.u1, a.dym, a.dy);
    return hipGetLastError();
}

}  // namespace cans
