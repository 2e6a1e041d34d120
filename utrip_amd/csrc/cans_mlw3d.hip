// cans_mlw3d.hip -- 3D modified Lax-Wendroff MHD step (cans3d/hdmlw/mlw_m.f)
// on gfx950, bit-exact FP64.
//
// Three kernels over SoA scratch in HBM (the 3D stencil's working set -- 8
// fields, 15 distinct cell fluxes, 18 corner fluxes -- does not fit a useful
// LDS tile at FP64):
//   cells    conserved variables, E = -v x B and the cell-centred fluxes
//            (mlw_m.f:70-192 flux loops), 23 values per cell;
//   corners  8-point half step (mlwhalf.f:38-68), half-step primitives and
//            the corner fluxes (mlw_m.f:193-315), 18 values per corner;
//   update   centre term of mlwhalf, mlwfull.f:34-55, mlwartv.f:36-48,
//            update and primitive recovery (mlw_m.f:316-453).
// The same kernels do the curvilinear engines (template G):
//   G = 1  cylindrical (r, phi, z), cans3d/hdmlw/mlw_m_c.f: phi fluxes divided
//          by the radius (x at cells, xm at corners) in slots of their own,
//          and the geometric sources of ro, ee, rx, ry, rz, bz;
//   G = 2  spherical (r, theta, phi), cans3d/hdmlw/mlw_m_s.f: the conserved
//          variables and fluxes carry the metric factors (r^2 sin, r sin, r,
//          ...), the momenta the geometric sources; the state is unscaled at
//          the corners and after the update.  sin, cos, tan of the
//          colatitudes come as per-row tables made on the host.
// Sources enter at cells into the corner half step (mlwsrch.f:39-46) and the
// centre term (mlwsrch.f:33-38), at corners into the update
// (mlwsrcf.f:33-46).
#include "cans_common.hpp"
#include "cans_clock.hpp"
#include "cans_roe.hpp"

namespace cans {
namespace {

constexpr double pi8i = 1.0 / kPi / 8.0;
constexpr double pi4i = 1.0 / kPi / 4.0;

// cell scratch slots
enum { C_RO, C_EE, C_RX, C_RY, C_RZ, C_BX, C_BY, C_BZ,
       C_EEX, C_EEY, C_EEZ, C_RXX, C_RXY, C_RXZ, C_RYX, C_RYY, C_RYZ, C_RZX, C_RZY, C_RZZ,
       C_EX, C_EY, C_EZ };
// corner scratch slots
enum { H_ROX, H_ROY, H_ROZ, H_EEX, H_EEY, H_EEZ, H_RXX, H_RXY, H_RXZ, H_RYX, H_RYY, H_RYZ,
       H_RZX, H_RZY, H_RZZ, H_EX, H_EY, H_EZ };

// flux of equation q (ro ee rx ry rz bx by bz) in direction d: scratch slot
// and sign; slot -1 is an identically zero flux.
__constant__ int kCellFlux[8][3][2] = {
    {{C_RX, 1}, {C_RY, 1}, {C_RZ, 1}},
    {{C_EEX, 1}, {C_EEY, 1}, {C_EEZ, 1}},
    {{C_RXX, 1}, {C_RXY, 1}, {C_RXZ, 1}},
    {{C_RYX, 1}, {C_RYY, 1}, {C_RYZ, 1}},
    {{C_RZX, 1}, {C_RZY, 1}, {C_RZZ, 1}},
    {{-1, 1}, {C_EZ, 1}, {C_EY, -1}},     // bx: 0, ez, -ey
    {{C_EZ, -1}, {-1, 1}, {C_EX, 1}},     // by: -ez, 0, ex
    {{C_EY, 1}, {C_EX, -1}, {-1, 1}},     // bz: ey, -ex, 0
};
// cylindrical cell slots (mlw_m_c.f:109-243): conserved, x fluxes, raw E,
// phi fluxes / x, z fluxes, sources
enum { Y_EEX = 8, Y_RXX, Y_RYX, Y_RZX, Y_EX, Y_EY, Y_EZ,
       Y_FRO, Y_FEE, Y_FRX, Y_FRY, Y_FRZ, Y_FBX, Y_FBZ,
       Y_EEZ, Y_RXZ, Y_RYZ, Y_RZZ,
       Y_SRO, Y_SEE, Y_SRX, Y_SRY, Y_SRZ, Y_SBZ, Y_NCELL };
static_assert(Y_NCELL == kMlw3dCylCellFields, "cylindrical cell slots");
// cylindrical corner slots (mlw_m_c.f:273-408)
enum { Z_ROX, Z_EEX, Z_RXX, Z_RYX, Z_RZX, Z_EX, Z_EY, Z_EZ,
       Z_FRO, Z_FEE, Z_FRX, Z_FRY, Z_FRZ, Z_FBX, Z_FBZ,
       Z_ROZ, Z_EEZ, Z_RXZ, Z_RYZ, Z_RZZ,
       Z_SRO, Z_SEE, Z_SRX, Z_SRY, Z_SRZ, Z_SBZ, Z_NCORN };
static_assert(Z_NCORN == kMlw3dCylCornerFields, "cylindrical corner slots");
__constant__ int kCellFluxC[8][3][2] = {
    {{C_RX, 1}, {Y_FRO, 1}, {C_RZ, 1}},
    {{Y_EEX, 1}, {Y_FEE, 1}, {Y_EEZ, 1}},
    {{Y_RXX, 1}, {Y_FRX, 1}, {Y_RXZ, 1}},
    {{Y_RYX, 1}, {Y_FRY, 1}, {Y_RYZ, 1}},
    {{Y_RZX, 1}, {Y_FRZ, 1}, {Y_RZZ, 1}},
    {{-1, 1}, {Y_FBX, 1}, {Y_EY, -1}},    // bx: 0, ez/x, -ey
    {{Y_EZ, -1}, {-1, 1}, {Y_EX, 1}},     // by: -ez, 0, ex
    {{Y_EY, 1}, {Y_FBZ, 1}, {-1, 1}},     // bz: ey, -ex/x, 0
};
__constant__ int kCornFluxC[8][3][2] = {
    {{Z_ROX, 1}, {Z_FRO, 1}, {Z_ROZ, 1}},
    {{Z_EEX, 1}, {Z_FEE, 1}, {Z_EEZ, 1}},
    {{Z_RXX, 1}, {Z_FRX, 1}, {Z_RXZ, 1}},
    {{Z_RYX, 1}, {Z_FRY, 1}, {Z_RYZ, 1}},
    {{Z_RZX, 1}, {Z_FRZ, 1}, {Z_RZZ, 1}},
    {{-1, 1}, {Z_FBX, 1}, {Z_EY, -1}},
    {{Z_EZ, -1}, {-1, 1}, {Z_EX, 1}},
    {{Z_EY, 1}, {Z_FBZ, 1}, {-1, 1}},
};
// source slot of each equation (bx, by: none)
__constant__ int kCellSrc[8] = {Y_SRO, Y_SEE, Y_SRX, Y_SRY, Y_SRZ, -1, -1, Y_SBZ};
__constant__ int kCornSrc[8] = {Z_SRO, Z_SEE, Z_SRX, Z_SRY, Z_SRZ, -1, -1, Z_SBZ};

// spherical cell slots (mlw_m_s.f:158-321): scaled conserved variables in
// 0..7, then the scaled x, y, z fluxes and the momentum sources
enum { P_XRO = 8, P_XEE, P_XRX, P_XRY, P_XRZ, P_XBY, P_XBZ,
       P_YRO, P_YEE, P_YRX, P_YRY, P_YRZ, P_YBX, P_YBZ,
       P_ZRO, P_ZEE, P_ZRX, P_ZRY, P_ZRZ, P_ZBX, P_ZBY,
       P_SRX, P_SRY, P_SRZ, P_NCELL };
static_assert(P_NCELL <= kMlw3dCylCellFields, "spherical cell slots");
// spherical corner slots (mlw_m_s.f:356-503)
enum { Q_XRO, Q_XEE, Q_XRX, Q_XRY, Q_XRZ, Q_XBY, Q_XBZ,
       Q_YRO, Q_YEE, Q_YRX, Q_YRY, Q_YRZ, Q_YBX, Q_YBZ,
       Q_ZRO, Q_ZEE, Q_ZRX, Q_ZRY, Q_ZRZ, Q_ZBX, Q_ZBY,
       Q_SRX, Q_SRY, Q_SRZ, Q_NCORN };
static_assert(Q_NCORN <= kMlw3dCylCornerFields, "spherical corner slots");
__constant__ int kCellFluxS[8][3][2] = {
    {{P_XRO, 1}, {P_YRO, 1}, {P_ZRO, 1}},
    {{P_XEE, 1}, {P_YEE, 1}, {P_ZEE, 1}},
    {{P_XRX, 1}, {P_YRX, 1}, {P_ZRX, 1}},
    {{P_XRY, 1}, {P_YRY, 1}, {P_ZRY, 1}},
    {{P_XRZ, 1}, {P_YRZ, 1}, {P_ZRZ, 1}},
    {{-1, 1}, {P_YBX, 1}, {P_ZBX, 1}},
    {{P_XBY, 1}, {-1, 1}, {P_ZBY, 1}},
    {{P_XBZ, 1}, {P_YBZ, 1}, {-1, 1}},
};
__constant__ int kCornFluxS[8][3][2] = {
    {{Q_XRO, 1}, {Q_YRO, 1}, {Q_ZRO, 1}},
    {{Q_XEE, 1}, {Q_YEE, 1}, {Q_ZEE, 1}},
    {{Q_XRX, 1}, {Q_YRX, 1}, {Q_ZRX, 1}},
    {{Q_XRY, 1}, {Q_YRY, 1}, {Q_ZRY, 1}},
    {{Q_XRZ, 1}, {Q_YRZ, 1}, {Q_ZRZ, 1}},
    {{-1, 1}, {Q_YBX, 1}, {Q_ZBX, 1}},
    {{Q_XBY, 1}, {-1, 1}, {Q_ZBY, 1}},
    {{Q_XBZ, 1}, {Q_YBZ, 1}, {-1, 1}},
};
__constant__ int kCellSrcS[8] = {-1, -1, P_SRX, P_SRY, P_SRZ, -1, -1, -1};
__constant__ int kCornSrcS[8] = {-1, -1, Q_SRX, Q_SRY, Q_SRZ, -1, -1, -1};

enum { G_CART = 0, G_CYL = 1, G_SPH = 2 };
template <int G> CANS_DEV int csrc(int q) { return G == G_CYL ? kCellSrc[q] : (G == G_SPH ? kCellSrcS[q] : -1); }
template <int G> CANS_DEV int hsrc(int q) { return G == G_CYL ? kCornSrc[q] : (G == G_SPH ? kCornSrcS[q] : -1); }

__constant__ int kCornFlux[8][3][2] = {
    {{H_ROX, 1}, {H_ROY, 1}, {H_ROZ, 1}},
    {{H_EEX, 1}, {H_EEY, 1}, {H_EEZ, 1}},
    {{H_RXX, 1}, {H_RXY, 1}, {H_RXZ, 1}},
    {{H_RYX, 1}, {H_RYY, 1}, {H_RYZ, 1}},
    {{H_RZX, 1}, {H_RZY, 1}, {H_RZZ, 1}},
    {{-1, 1}, {H_EZ, 1}, {H_EY, -1}},
    {{H_EZ, -1}, {-1, 1}, {H_EX, 1}},
    {{H_EY, 1}, {H_EX, -1}, {-1, 1}},
};

// The three kernels take their blocks in the XCD-contiguous order of
// xcd_first_thread() (cans_common.hpp).
CANS_DEV bool step_dt3(const Mlw3dArgs &a, double &dt) {
    if (!a.clk) { dt = a.dt; return true; }
    const int d = (a.step + 1) & 1;
    if (a.clk->stopped[d]) return false;
    dt = a.clk->dt_last[d];
    return true;
}

template <int G>
__global__ __launch_bounds__(256) void mlw3d_cells(Mlw3dArgs a) {
    double dt;
    if (!step_dt3(a, dt)) return;
    const size_t N = (size_t)a.ix * a.jx * a.kx;
    const size_t p = xcd_first_thread() + threadIdx.x;
    if (p >= N) return;
    const double gm = a.gm;
    const double ro = a.prim[0][p], pr = a.prim[1][p], vx = a.prim[2][p], vy = a.prim[3][p], vz = a.prim[4][p];
    const double bx = a.prim[5][p], by = a.prim[6][p], bz = a.prim[7][p];
    double *c = a.cell;
#define W(slot, v) c[(size_t)(slot) * N + p] = (v)
    const double vv = vx * vx + vy * vy + vz * vz;
    const double bb = bx * bx + by * by + bz * bz;
    if (G == G_SPH) {   // mlw_m_s.f:137-321, r = x(i), st = sin(y(j))
        const size_t i0 = p % (size_t)a.ix, j0 = (p / (size_t)a.ix) % (size_t)a.jx;
        const double r = a.x[i0];
        const double st = a.trig[j0], ct = a.trig[a.jx + j0], tt = a.trig[2 * a.jx + j0];
        const double r2 = r * r;
        const double ee = pr / (gm - 1.0) + 0.5 * ro * vv + pi8i * bb;
        W(0, ro * r2 * st);
        W(1, ee * r2 * st);
        W(2, ro * vx * r2 * st);
        W(3, ro * vy * r * st);
        W(4, ro * vz * r);
        W(5, bx * r2 * st);
        W(6, by * r * st);
        W(7, bz * r);
        const double ex = -vy * bz + vz * by;
        const double ey = -vz * bx + vx * bz;
        const double ez = -vx * by + vy * bx;
        W(P_XRO, ro * vx * r2 * st);
        W(P_YRO, ro * vy * r * st);
        W(P_ZRO, ro * vz * r);
        const double ep = pr * gm / (gm - 1.0) + 0.5 * ro * vv;
        const double px = (bz * ey - by * ez) * pi4i;
        const double py = (bx * ez - bz * ex) * pi4i;
        const double pz = (by * ex - bx * ey) * pi4i;
        W(P_XEE, (ep * vx + px) * r2 * st);
        W(P_YEE, (ep * vy + py) * r * st);
        W(P_ZEE, (ep * vz + pz) * r);
        const double rxx = ro * (vx * vx) + pr + pi8i * (by * by + bz * bz - bx * bx);
        const double rxy = ro * vx * vy - pi4i * bx * by;
        const double rxz = ro * vx * vz - pi4i * bx * bz;
        W(P_XRX, rxx * r2 * st);
        W(P_YRX, rxy * r * st);
        W(P_ZRX, rxz * r);
        W(P_SRX, (ro * (vy * vy + vz * vz) + 2.0 * pr + pi4i * (bx * bx)) * r * st);
        const double ryx = ro * vy * vx - pi4i * by * bx;
        const double ryy = ro * (vy * vy) + pr + pi8i * (bx * bx + bz * bz - by * by);
        const double ryz = ro * vy * vz - pi4i * by * bz;
        const double rzz = ro * (vz * vz) + pr + pi8i * (bx * bx + by * by - bz * bz);
        W(P_XRY, ryx * r * st);
        W(P_YRY, ryy * st);
        W(P_ZRY, ryz);
        W(P_SRY, rxy * (-2.0) * st + rzz * ct);
        const double rzx = ro * vz * vx - pi4i * bz * bx;
        const double rzy = ro * vz * vy - pi4i * bz * by;
        W(P_XRZ, rzx * r);
        W(P_YRZ, rzy);
        W(P_ZRZ, rzz / st);
        W(P_SRZ, rxz * (-2.0) + ryz * (-2.0) / tt);
        W(P_YBX, ez * r * st);
        W(P_ZBX, -ey * r);
        W(P_XBY, -ez * r * st);
        W(P_ZBY, ex);
        W(P_XBZ, ey * r);
        W(P_YBZ, -ex);
        return;
    }
    W(C_RO, ro);
    W(C_EE, pr / (gm - 1.0) + 0.5 * ro * vv + pi8i * bb);
    W(C_RX, ro * vx);
    W(C_RY, ro * vy);
    W(C_RZ, ro * vz);
    W(C_BX, bx);
    W(C_BY, by);
    W(C_BZ, bz);
    const double ex = -vy * bz + vz * by;
    const double ey = -vz * bx + vx * bz;
    const double ez = -vx * by + vy * bx;
    const double ep = pr * gm / (gm - 1.0) + 0.5 * ro * vv;
    if (G == G_CYL) {   // mlw_m_c.f:109-243, r = x(i)
        const double r = a.x[p % (size_t)a.ix];
        const double fro = ro * vx;
        W(Y_FRO, ro * vy / r);
        W(Y_SRO, -(fro / r));
        const double fee = ep * vx + (bz * ey - by * ez) * pi4i;
        W(Y_EEX, fee);
        W(Y_FEE, (ep * vy + (bx * ez - bz * ex) * pi4i) / r);
        W(Y_EEZ, ep * vz + (by * ex - bx * ey) * pi4i);
        W(Y_SEE, -(fee / r));
        W(Y_RXX, ro * (vx * vx) + pr + pi8i * (by * by + bz * bz - bx * bx));
        W(Y_FRX, (ro * vx * vy - pi4i * bx * by) / r);
        W(Y_RXZ, ro * vx * vz - pi4i * bx * bz);
        W(Y_SRX, -(ro / r * (vx * vx - vy * vy)) + pi4i / r * (bx * bx - by * by));
        W(Y_RYX, ro * vy * vx - pi4i * by * bx);
        W(Y_FRY, (ro * (vy * vy) + pr + pi8i * (bx * bx + bz * bz - by * by)) / r);
        W(Y_RYZ, ro * vy * vz - pi4i * by * bz);
        W(Y_SRY, -(2.0 * ro / r * vx * vy) + 2.0 * pi4i / r * bx * by);
        const double frz = ro * vz * vx - pi4i * bz * bx;
        W(Y_RZX, frz);
        W(Y_FRZ, (ro * vz * vy - pi4i * bz * by) / r);
        W(Y_RZZ, ro * (vz * vz) + pr + pi8i * (bx * bx + by * by - bz * bz));
        W(Y_SRZ, -(frz / r));
        W(Y_EX, ex);
        W(Y_EY, ey);
        W(Y_EZ, ez);
        W(Y_FBX, ez / r);
        W(Y_FBZ, -ex / r);
        W(Y_SBZ, -(ey / r));
        return;
    }
    W(C_EEX, ep * vx + (bz * ey - by * ez) * pi4i);
    W(C_EEY, ep * vy + (bx * ez - bz * ex) * pi4i);
    W(C_EEZ, ep * vz + (by * ex - bx * ey) * pi4i);
    W(C_RXX, ro * (vx * vx) + pr + pi8i * (by * by + bz * bz - bx * bx));
    W(C_RXY, ro * vx * vy - pi4i * bx * by);
    W(C_RXZ, ro * vx * vz - pi4i * bx * bz);
    W(C_RYX, ro * vy * vx - pi4i * by * bx);
    W(C_RYY, ro * (vy * vy) + pr + pi8i * (bx * bx + bz * bz - by * by));
    W(C_RYZ, ro * vy * vz - pi4i * by * bz);
    W(C_RZX, ro * vz * vx - pi4i * bz * bx);
    W(C_RZY, ro * vz * vy - pi4i * bz * by);
    W(C_RZZ, ro * (vz * vz) + pr + pi8i * (bx * bx + by * by - bz * bz));
    W(C_EX, ex);
    W(C_EY, ey);
    W(C_EZ, ez);
#undef W
}

template <int G>
CANS_DEV double cflux(const double *c, size_t N, int q, int d, size_t p) {
    const int(*t)[3][2] = G == G_CYL ? kCellFluxC : (G == G_SPH ? kCellFluxS : kCellFlux);
    const int slot = t[q][d][0];
    if (slot < 0) return 0.0;
    const double v = c[(size_t)slot * N + p];
    return t[q][d][1] < 0 ? -v : v;
}
template <int G>
CANS_DEV double hflux(const double *h, size_t N, int q, int d, size_t p) {
    const int(*t)[3][2] = G == G_CYL ? kCornFluxC : (G == G_SPH ? kCornFluxS : kCornFlux);
    const int slot = t[q][d][0];
    if (slot < 0) return 0.0;
    const double v = h[(size_t)slot * N + p];
    return t[q][d][1] < 0 ? -v : v;
}

// corner (i,j,k) between cells i..i+1, j..j+1, k..k+1, i,j,k in [1,n-1]
template <int G>
__global__ __launch_bounds__(256) void mlw3d_corners(Mlw3dArgs a) {
    double dt;
    if (!step_dt3(a, dt)) return;
    const int ix = a.ix, jx = a.jx, kx = a.kx;
    const int ci = ix - 1, cj = jx - 1, ck = kx - 1;
    const size_t t = xcd_first_thread() + threadIdx.x;
    if (t >= (size_t)ci * cj * ck) return;
    const int i = 1 + (int)(t % ci);
    const int j = 1 + (int)((t / ci) % cj);
    const int k = 1 + (int)(t / ((size_t)ci * cj));
    const size_t N = (size_t)ix * jx * kx;
    const size_t sy = ix, sz = (size_t)ix * jx;
    const size_t p000 = (size_t)(i - 1) + sy * (j - 1) + sz * (k - 1);
    const size_t p100 = p000 + 1, p010 = p000 + sy, p110 = p010 + 1;
    const size_t p001 = p000 + sz, p101 = p001 + 1, p011 = p001 + sy, p111 = p011 + 1;
    const double *c = a.cell;
    const double dxim = a.mx.dim[i - 1], dyim = a.my.dim[j - 1], dzim = a.mz.dim[k - 1];
    double un[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const double *u = c + (size_t)q * N;    // C_RO..C_BZ are slots 0..7
        const double uh = (u[p100] + u[p000] + u[p110] + u[p010] + u[p101] + u[p001] + u[p111] + u[p011]) / 8.0;
        const double dfdx = (cflux<G>(c, N, q, 0, p100) - cflux<G>(c, N, q, 0, p000) + cflux<G>(c, N, q, 0, p110)
                             - cflux<G>(c, N, q, 0, p010) + cflux<G>(c, N, q, 0, p101) - cflux<G>(c, N, q, 0, p001)
                             + cflux<G>(c, N, q, 0, p111) - cflux<G>(c, N, q, 0, p011)) / 4.0 * dxim;
        const double dfdy = (cflux<G>(c, N, q, 1, p010) - cflux<G>(c, N, q, 1, p000) + cflux<G>(c, N, q, 1, p110)
                             - cflux<G>(c, N, q, 1, p100) + cflux<G>(c, N, q, 1, p011) - cflux<G>(c, N, q, 1, p001)
                             + cflux<G>(c, N, q, 1, p111) - cflux<G>(c, N, q, 1, p101)) / 4.0 * dyim;
        const double dfdz = (cflux<G>(c, N, q, 2, p001) - cflux<G>(c, N, q, 2, p000) + cflux<G>(c, N, q, 2, p101)
                             - cflux<G>(c, N, q, 2, p100) + cflux<G>(c, N, q, 2, p011) - cflux<G>(c, N, q, 2, p010)
                             + cflux<G>(c, N, q, 2, p111) - cflux<G>(c, N, q, 2, p110)) / 4.0 * dzim;
        un[q] = uh - dt * dfdx - dt * dfdy - dt * dfdz;
        if (csrc<G>(q) >= 0) {   // mlwsrch.f:39-46
            const double *sc = c + (size_t)csrc<G>(q) * N;
            const double sh = (sc[p100] + sc[p000] + sc[p110] + sc[p010] + sc[p101] + sc[p001] + sc[p111] + sc[p011])
                              / 8.0;
            un[q] = un[q] + dt * sh;
        }
    }
    const double gm = a.gm;
    double roh, eeh, vxh, vyh, vzh, bxh, byh, bzh;
    double r = 0.0, r2 = 0.0, st = 0.0;
    if (G == G_SPH) {   // mlw_m_s.f:325-343, r = xm(i), st = sin(ym(j))
        r = a.xm[i - 1];
        st = a.trig[3 * a.jx + j - 1];
        r2 = r * r;
        roh = un[0] / r2 / st;
        eeh = un[1] / r2 / st;
        vxh = un[2] / roh / r2 / st;
        vyh = un[3] / roh / r / st;
        vzh = un[4] / roh / r;
        bxh = un[5] / r2 / st;
        byh = un[6] / r / st;
        bzh = un[7] / r;
    } else {
        roh = un[0]; eeh = un[1]; bxh = un[5]; byh = un[6]; bzh = un[7];
        vxh = un[2] / roh;
        vyh = un[3] / roh;
        vzh = un[4] / roh;
    }
    const double vv = vxh * vxh + vyh * vyh + vzh * vzh;
    const double bb = bxh * bxh + byh * byh + bzh * bzh;
    const double prh = (gm - 1.0) * (eeh - 0.5 * roh * vv - pi8i * bb);
    const double exh = -vyh * bzh + vzh * byh;
    const double eyh = -vzh * bxh + vxh * bzh;
    const double ezh = -vxh * byh + vyh * bxh;
    const double ep = prh * gm / (gm - 1.0) + 0.5 * roh * vv;
    double *h = a.corner;
    const size_t pc = p000;   // corner (i,j,k) stored at the index of cell (i,j,k)
#define W(slot, v) h[(size_t)(slot) * N + pc] = (v)
    if (G == G_SPH) {   // mlw_m_s.f:356-503
        const double ct = a.trig[4 * a.jx + j - 1], tt = a.trig[5 * a.jx + j - 1];
        W(Q_XRO, roh * vxh * r2 * st);
        W(Q_YRO, roh * vyh * r * st);
        W(Q_ZRO, roh * vzh * r);
        const double px = (bzh * eyh - byh * ezh) * pi4i;
        const double py = (bxh * ezh - bzh * exh) * pi4i;
        const double pz = (byh * exh - bxh * eyh) * pi4i;
        W(Q_XEE, (ep * vxh + px) * r2 * st);
        W(Q_YEE, (ep * vyh + py) * r * st);
        W(Q_ZEE, (ep * vzh + pz) * r);
        const double rxx = roh * (vxh * vxh) + prh + pi8i * (byh * byh + bzh * bzh - bxh * bxh);
        const double rxy = roh * vxh * vyh - pi4i * bxh * byh;
        const double rxz = roh * vxh * vzh - pi4i * bxh * bzh;
        W(Q_XRX, rxx * r2 * st);
        W(Q_YRX, rxy * r * st);
        W(Q_ZRX, rxz * r);
        W(Q_SRX, (roh * (vyh * vyh + vzh * vzh) + 2.0 * prh + pi4i * (bxh * bxh)) * r * st);
        const double ryx = roh * vyh * vxh - pi4i * byh * bxh;
        const double ryy = roh * (vyh * vyh) + prh + pi8i * (bxh * bxh + bzh * bzh - byh * byh);
        const double ryz = roh * vyh * vzh - pi4i * byh * bzh;
        const double rzz = roh * (vzh * vzh) + prh + pi8i * (bxh * bxh + byh * byh - bzh * bzh);
        W(Q_XRY, ryx * r * st);
        W(Q_YRY, ryy * st);
        W(Q_ZRY, ryz);
        W(Q_SRY, rxy * (-2.0) * st + rzz * ct);
        const double rzx = roh * vzh * vxh - pi4i * bzh * bxh;
        const double rzy = roh * vzh * vyh - pi4i * bzh * byh;
        W(Q_XRZ, rzx * r);
        W(Q_YRZ, rzy);
        W(Q_ZRZ, rzz / st);
        W(Q_SRZ, rxz * (-2.0) + ryz * (-2.0) / tt);
        W(Q_YBX, ezh * r * st);
        W(Q_ZBX, -eyh * r);
        W(Q_XBY, -ezh * r * st);
        W(Q_ZBY, exh);
        W(Q_XBZ, eyh * r);
        W(Q_YBZ, -exh);
        return;
    }
    if (G == G_CYL) {   // mlw_m_c.f:273-408, r = xm(i)
        const double r = a.xm[i - 1];
        const double fro = roh * vxh;
        W(Z_ROX, fro);
        W(Z_FRO, roh * vyh / r);
        W(Z_ROZ, roh * vzh);
        W(Z_SRO, -(fro / r));
        const double fee = ep * vxh + (bzh * eyh - byh * ezh) * pi4i;
        W(Z_EEX, fee);
        W(Z_FEE, (ep * vyh + (bxh * ezh - bzh * exh) * pi4i) / r);
        W(Z_EEZ, ep * vzh + (byh * exh - bxh * eyh) * pi4i);
        W(Z_SEE, -(fee / r));
        W(Z_RXX, roh * (vxh * vxh) + prh + pi8i * (byh * byh + bzh * bzh - bxh * bxh));
        W(Z_FRX, (roh * vxh * vyh - pi4i * bxh * byh) / r);
        W(Z_RXZ, roh * vxh * vzh - pi4i * bxh * bzh);
        W(Z_SRX, -(roh / r * (vxh * vxh - vyh * vyh)) + pi4i / r * (bxh * bxh - byh * byh));
        W(Z_RYX, roh * vyh * vxh - pi4i * byh * bxh);
        W(Z_FRY, (roh * (vyh * vyh) + prh + pi8i * (bxh * bxh + bzh * bzh - byh * byh)) / r);
        W(Z_RYZ, roh * vyh * vzh - pi4i * byh * bzh);
        W(Z_SRY, -(2.0 * roh / r * vxh * vyh) + 2.0 * pi4i / r * bxh * byh);
        const double frz = roh * vzh * vxh - pi4i * bzh * bxh;
        W(Z_RZX, frz);
        W(Z_FRZ, (roh * vzh * vyh - pi4i * bzh * byh) / r);
        W(Z_RZZ, roh * (vzh * vzh) + prh + pi8i * (bxh * bxh + byh * byh - bzh * bzh));
        W(Z_SRZ, -(frz / r));
        W(Z_EX, exh);
        W(Z_EY, eyh);
        W(Z_EZ, ezh);
        W(Z_FBX, ezh / r);
        W(Z_FBZ, -exh / r);
        W(Z_SBZ, -(eyh / r));
        return;
    }
    W(H_ROX, roh * vxh);
    W(H_ROY, roh * vyh);
    W(H_ROZ, roh * vzh);
    W(H_EEX, ep * vxh + (bzh * eyh - byh * ezh) * pi4i);
    W(H_EEY, ep * vyh + (bxh * ezh - bzh * exh) * pi4i);
    W(H_EEZ, ep * vzh + (byh * exh - bxh * eyh) * pi4i);
    W(H_RXX, roh * (vxh * vxh) + prh + pi8i * (byh * byh + bzh * bzh - bxh * bxh));
    W(H_RXY, roh * vxh * vyh - pi4i * bxh * byh);
    W(H_RXZ, roh * vxh * vzh - pi4i * bxh * bzh);
    W(H_RYX, roh * vyh * vxh - pi4i * byh * bxh);
    W(H_RYY, roh * (vyh * vyh) + prh + pi8i * (bxh * bxh + bzh * bzh - byh * byh));
    W(H_RYZ, roh * vyh * vzh - pi4i * byh * bzh);
    W(H_RZX, roh * vzh * vxh - pi4i * bzh * bxh);
    W(H_RZY, roh * vzh * vyh - pi4i * bzh * byh);
    W(H_RZZ, roh * (vzh * vzh) + prh + pi8i * (bxh * bxh + byh * byh - bzh * bzh));
    W(H_EX, exh);
    W(H_EY, eyh);
    W(H_EZ, ezh);
#undef W
}

// interior cell (i,j,k) in [2,n-1]^3
template <int G>
__global__ __launch_bounds__(256) void mlw3d_update(Mlw3dArgs a) {
    double dt;
    if (!step_dt3(a, dt)) return;
    const int ix = a.ix, jx = a.jx, kx = a.kx;
    const int ni = ix - 2, nj = jx - 2, nk = kx - 2;
    const size_t t = xcd_first_thread() + threadIdx.x;
    if (t >= (size_t)ni * nj * nk) return;
    const int i = 2 + (int)(t % ni);
    const int j = 2 + (int)((t / ni) % nj);
    const int k = 2 + (int)(t / ((size_t)ni * nj));
    const size_t N = (size_t)ix * jx * kx;
    const size_t sx = 1, sy = ix, sz = (size_t)ix * jx;
    const size_t p = (size_t)(i - 1) + sy * (j - 1) + sz * (k - 1);
    const double *c = a.cell, *h = a.corner;
    const double dxi = a.mx.di[i - 1], dyi = a.my.di[j - 1], dzi = a.mz.di[k - 1];
    const double dxim = a.mx.dim[i - 1], dyim = a.my.dim[j - 1], dzim = a.mz.dim[k - 1];
    const double dxim1 = a.mx.dim[i - 2], dyim1 = a.my.dim[j - 2], dzim1 = a.mz.dim[k - 2];
    const double ux0 = a.mx.u0[i - 1], ux1 = a.mx.u1[i - 1];
    const double uy0 = a.my.u0[j - 1], uy1 = a.my.u1[j - 1];
    const double uz0 = a.mz.u0[k - 1], uz1 = a.mz.u1[k - 1];
    // corners around the cell: (i-1|i, j-1|j, k-1|k)
    const size_t c000 = p - sx - sy - sz;
    const size_t c100 = c000 + sx, c010 = c000 + sy, c110 = c010 + sx;
    const size_t c001 = c000 + sz, c101 = c001 + sx, c011 = c001 + sy, c111 = c011 + sx;
    // artificial viscosity coefficients (mlw_m.f:316-337)
    const double zero = 0.0;
    const double *vxp = a.prim[2], *vyp = a.prim[3], *vzp = a.prim[4];
    const double qxc = a.qav * a.dxm[i - 1] * fmaxF(zero, __builtin_fabs(vxp[p + sx] - vxp[p]) - kR4_1em4);
    const double qxw = a.qav * a.dxm[i - 2] * fmaxF(zero, __builtin_fabs(vxp[p] - vxp[p - sx]) - kR4_1em4);
    const double qyc = a.qav * a.dym[j - 1] * fmaxF(zero, __builtin_fabs(vyp[p + sy] - vyp[p]) - kR4_1em4);
    const double qys = a.qav * a.dym[j - 2] * fmaxF(zero, __builtin_fabs(vyp[p] - vyp[p - sy]) - kR4_1em4);
    const double qzc = a.qav * a.dzm[k - 1] * fmaxF(zero, __builtin_fabs(vzp[p + sz] - vzp[p]) - kR4_1em4);
    const double qzb = a.qav * a.dzm[k - 2] * fmaxF(zero, __builtin_fabs(vzp[p] - vzp[p - sz]) - kR4_1em4);
    double U[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        double du = 0.0;
        // mlwhalf centre term
        du = du - 0.5 * dt * (0.5 * dxi * (cflux<G>(c, N, q, 0, p + sx) - cflux<G>(c, N, q, 0, p - sx)))
                - 0.5 * dt * (0.5 * dyi * (cflux<G>(c, N, q, 1, p + sy) - cflux<G>(c, N, q, 1, p - sy)))
                - 0.5 * dt * (0.5 * dzi * (cflux<G>(c, N, q, 2, p + sz) - cflux<G>(c, N, q, 2, p - sz)));
        if (csrc<G>(q) >= 0) du = du + 0.5 * dt * c[(size_t)csrc<G>(q) * N + p];   // mlwsrch.f:33-38
        // mlwfull
        const double dfdx = dxi * (uz0 * (uy0 * (hflux<G>(h, N, q, 0, c100) - hflux<G>(h, N, q, 0, c000))
                                          + uy1 * (hflux<G>(h, N, q, 0, c110) - hflux<G>(h, N, q, 0, c010)))
                                   + uz1 * (uy0 * (hflux<G>(h, N, q, 0, c101) - hflux<G>(h, N, q, 0, c001))
                                            + uy1 * (hflux<G>(h, N, q, 0, c111) - hflux<G>(h, N, q, 0, c011))));
        const double dfdy = dyi * (ux0 * (uz0 * (hflux<G>(h, N, q, 1, c010) - hflux<G>(h, N, q, 1, c000))
                                          + uz1 * (hflux<G>(h, N, q, 1, c011) - hflux<G>(h, N, q, 1, c001)))
                                   + ux1 * (uz0 * (hflux<G>(h, N, q, 1, c110) - hflux<G>(h, N, q, 1, c100))
                                            + uz1 * (hflux<G>(h, N, q, 1, c111) - hflux<G>(h, N, q, 1, c101))));
        const double dfdz = dzi * (uy0 * (ux0 * (hflux<G>(h, N, q, 2, c001) - hflux<G>(h, N, q, 2, c000))
                                          + ux1 * (hflux<G>(h, N, q, 2, c101) - hflux<G>(h, N, q, 2, c100)))
                                   + uy1 * (ux0 * (hflux<G>(h, N, q, 2, c011) - hflux<G>(h, N, q, 2, c010))
                                            + ux1 * (hflux<G>(h, N, q, 2, c111) - hflux<G>(h, N, q, 2, c110))));
        du = du - 0.5 * dt * dfdx - 0.5 * dt * dfdy - 0.5 * dt * dfdz;
        if (hsrc<G>(q) >= 0) {   // mlwsrcf.f:33-46
            const double *sh_ = h + (size_t)hsrc<G>(q) * N;
            const double sh = uz0 * (uy0 * (ux0 * sh_[c000] + ux1 * sh_[c100]) + uy1 * (ux0 * sh_[c010] + ux1 * sh_[c110]))
                              + uz1 * (uy0 * (ux0 * sh_[c001] + ux1 * sh_[c101])
                                       + uy1 * (ux0 * sh_[c011] + ux1 * sh_[c111]));
            du = du + 0.5 * dt * sh;
        }
        // mlwartv
        const double *u = c + (size_t)q * N;
        const double uc = u[p];
        du = du + dt * (dxi * (qxc * dxim * (u[p + sx] - uc) - qxw * dxim1 * (uc - u[p - sx])))
                + dt * (dyi * (qyc * dyim * (u[p + sy] - uc) - qys * dyim1 * (uc - u[p - sy])))
                + dt * (dzi * (qzc * dzim * (u[p + sz] - uc) - qzb * dzim1 * (uc - u[p - sz])));
        U[q] = uc + du;
    }
    const double gm = a.gm;
    if (G == G_SPH) {   // mlw_m_s.f:570-586, r = x(i), st = sin(y(j))
        const double r = a.x[i - 1], st = a.trig[j - 1], r2 = r * r;
        const double ro = U[0] / r2 / st;
        const double ee = U[1] / r2 / st;
        const double vx = U[2] / ro / r2 / st;
        const double vy = U[3] / ro / r / st;
        const double vz = U[4] / ro / r;
        const double bx = U[5] / r2 / st;
        const double by = U[6] / r / st;
        const double bz = U[7] / r;
        const double vv = vx * vx + vy * vy + vz * vz;
        const double bb = bx * bx + by * by + bz * bz;
        a.out[0][p] = ro;
        a.out[1][p] = (gm - 1.0) * (ee - 0.5 * ro * vv - pi8i * bb);
        a.out[2][p] = vx;
        a.out[3][p] = vy;
        a.out[4][p] = vz;
        a.out[5][p] = bx;
        a.out[6][p] = by;
        a.out[7][p] = bz;
        return;
    }
    const double ro = U[0];
    const double vx = U[2] / ro;
    const double vy = U[3] / ro;
    const double vz = U[4] / ro;
    const double vv = vx * vx + vy * vy + vz * vz;
    const double bb = U[5] * U[5] + U[6] * U[6] + U[7] * U[7];
    a.out[0][p] = ro;
    a.out[1][p] = (gm - 1.0) * (U[1] - 0.5 * ro * vv - pi8i * bb);
    a.out[2][p] = vx;
    a.out[3][p] = vy;
    a.out[4][p] = vz;
    a.out[5][p] = U[5];
    a.out[6][p] = U[6];
    a.out[7][p] = U[7];
}

}  // namespace

hipError_t launch_mlw3d(const Mlw3dArgs &a, hipStream_t s) {
    if (a.ix < 3 || a.jx < 3 || a.kx < 3) return hipSuccess;
    const size_t N = (size_t)a.ix * a.jx * a.kx;
    const size_t nc = (size_t)(a.ix - 1) * (a.jx - 1) * (a.kx - 1);
    const size_t nu = (size_t)(a.ix - 2) * (a.jx - 2) * (a.kx - 2);
    if (a.x && !a.xm) return hipErrorInvalidValue;
    if (a.trig && !a.x) return hipErrorInvalidValue;
    // (a periodic-box call, per_margin > 0, needs the marching kernel: it
    // carries the step's clock and cfl_m)
    if (!a.x && !a.trig && (a.per_margin > 0 || mlw3d_path() == 1)) return launch_mlw3d_march(a, s);
    const dim3 gc(xcd_grid(N)), gk(xcd_grid(nc)), gu(xcd_grid(nu));
    if (a.trig) {
        hipLaunchKernelGGL(mlw3d_cells<G_SPH>, gc, dim3(256), 0, s, a);
        hipLaunchKernelGGL(mlw3d_corners<G_SPH>, gk, dim3(256), 0, s, a);
        hipLaunchKernelGGL(mlw3d_update<G_SPH>, gu, dim3(256), 0, s, a);
    } else if (a.x) {
        hipLaunchKernelGGL(mlw3d_cells<G_CYL>, gc, dim3(256), 0, s, a);
        hipLaunchKernelGGL(mlw3d_corners<G_CYL>, gk, dim3(256), 0, s, a);
        hipLaunchKernelGGL(mlw3d_update<G_CYL>, gu, dim3(256), 0, s, a);
    } else {
        hipLaunchKernelGGL(mlw3d_cells<G_CART>, gc, dim3(256), 0, s, a);
        hipLaunchKernelGGL(mlw3d_corners<G_CART>, gk, dim3(256), 0, s, a);
        hipLaunchKernelGGL(mlw3d_update<G_CART>, gu, dim3(256), 0, s, a);
    }
    return hipGetLastError();
}

}  // namespace cans
