// cans_mlw2d_s.hip -- 2D spherical (r, theta) modified Lax-Wendroff MHD step
// with the vector potential Az (cans2d/hdmlw/mlw_m_s.f), bit-exact FP64.
//
// Three kernels over SoA scratch, as the 3D engines:
//   cells    metric-scaled conserved variables (ro ee rx by x^2 sin, ry by
//            x sin), the scaled cell fluxes and the sources of rx, ry, Az
//            (mlw_m_s.f:136-231);
//   corners  4-point half step of the scaled variables, mlwsrch's corner
//            term, unscaling by xm, sin(ym), half-step primitives and the
//            scaled corner fluxes and sources (mlwhalf.f, mlwsrch.f,
//            mlw_m_s.f:234-340);
//   update   centre term of mlwhalf, the sources at cells (mlwsrch) and
//            corners (mlwsrcf), mlwfull, mlwartv on the scaled variables,
//            update and unscaling (mlw_m_s.f:344-396).
// sin and cos of y and ym come as tables made on the host with the C
// library.  The corner x flux of ry carries sin(ym) without the radius
// (mlw_m_s.f:305), as in the reference.
#include "cans_common.hpp"
#include "cans_clock.hpp"
#include "cans_kernels.hpp"

namespace cans {
namespace {

constexpr double pi8i = 1.0 / kPi / 8.0;   // mlw_m_s.f:94-95, pi = acos(-1.0d0)
constexpr double pi4i = 1.0 / kPi / 4.0;

// scaled conserved variables ro ee rx ry bx by in cell slots 0..5
enum { S_XRO = 6, S_XEE, S_XRX, S_XRY, S_XBY, S_YRO, S_YEE, S_YRX, S_YRY, S_YBX, S_SRX, S_SRY, S_SAZ, S_NCELL };
enum { K_XRO, K_XEE, K_XRX, K_XRY, K_XBY, K_YRO, K_YEE, K_YRX, K_YRY, K_YBX, K_SRX, K_SRY, K_SAZ, K_NCORN };
static_assert(S_NCELL == kMlw2sCellFields && K_NCORN == kMlw2sCornerFields, "2D spherical slots");
// flux slot of equation q (ro ee rx ry bx by) in direction d; -1: zero
__constant__ int kFx[6] = {S_XRO, S_XEE, S_XRX, S_XRY, -1, S_XBY};
__constant__ int kFy[6] = {S_YRO, S_YEE, S_YRX, S_YRY, S_YBX, -1};
__constant__ int kHx[6] = {K_XRO, K_XEE, K_XRX, K_XRY, -1, K_XBY};
__constant__ int kHy[6] = {K_YRO, K_YEE, K_YRX, K_YRY, K_YBX, -1};

CANS_DEV bool step_dt2s(const Mlw2sArgs &a, double &dt) {
    if (!a.clk) { dt = a.dt; return true; }
    const int d = (a.step + 1) & 1;
    if (a.clk->stopped[d]) return false;
    dt = a.clk->dt_last[d];
    return true;
}

CANS_DEV double fl(const double *c, size_t N, int slot, size_t p) {
    return slot < 0 ? 0.0 : c[(size_t)slot * N + p];
}

__global__ __launch_bounds__(256) void mlw2s_cells(Mlw2sArgs a) {
    double dt;
    if (!step_dt2s(a, dt)) return;
    const size_t N = (size_t)a.ix * a.jx;
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const int i0 = (int)(p % a.ix), j0 = (int)(p / a.ix);
    const double gm = a.gm;
    const double ro = a.prim[0][p], pr = a.prim[1][p], vx = a.prim[2][p], vy = a.prim[3][p];
    const double bx = a.prim[4][p], by = a.prim[5][p];
    const double r = a.x[i0], st = a.trig[j0], ct = a.trig[a.jx + j0];
    const double r2 = r * r;
    double *c = a.cell;
#define W(slot, v) c[(size_t)(slot) * N + p] = (v)
    const double vv = vx * vx + vy * vy;
    const double bb = bx * bx + by * by;
    const double ee = pr / (gm - 1.0) + 0.5 * ro * vv + pi8i * bb;   // :113-119
    const double ez = -vx * by + vy * bx;
    W(0, ro * r2 * st);                                               // :136-145
    W(1, ee * r2 * st);
    W(2, ro * vx * r2 * st);
    W(3, ro * vy * r * st);
    W(4, bx * r2 * st);
    W(5, by * r * st);
    W(S_XRO, ro * vx * r2 * st);                                      // :148-157
    W(S_YRO, ro * vy * r * st);
    const double ep = pr * gm / (gm - 1.0) + 0.5 * ro * vv;          // :160-170
    W(S_XEE, (ep * vx - by * ez * pi4i) * r2 * st);
    W(S_YEE, (ep * vy + bx * ez * pi4i) * r * st);
    const double rxy = ro * vx * vy - pi4i * bx * by;                 // :173-184
    W(S_XRX, (ro * (vx * vx) + pr + pi8i * (by * by - bx * bx)) * r2 * st);
    W(S_YRX, rxy * r * st);
    W(S_SRX, (ro * (vy * vy) + 2.0 * pr + pi4i * (bx * bx)) * r * st);
    W(S_XRY, (ro * vy * vx - pi4i * by * bx) * r * st);              // :188-201
    W(S_YRY, (ro * (vy * vy) + pr + pi8i * (bx * bx - by * by)) * st);
    W(S_SRY, rxy * (-2.0) * st + (pr + pi8i * (bx * bx + by * by)) * ct);
    W(S_YBX, ez * r * st);                                            // :206-212
    W(S_XBY, -ez * r * st);                                           // :216-222
    W(S_SAZ, -ez);                                                    // :226-231
#undef W
}

// corner (i, j) between cells i..i+1, j..j+1 (i, j in [1, n-1]), stored at
// the index of cell (i, j)
__global__ __launch_bounds__(256) void mlw2s_corners(Mlw2sArgs a) {
    double dt;
    if (!step_dt2s(a, dt)) return;
    const int ix = a.ix, jx = a.jx, ci = ix - 1;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)ci * (jx - 1)) return;
    const int i = 1 + (int)(t % ci), j = 1 + (int)(t / ci);
    const size_t N = (size_t)ix * jx;
    const size_t p00 = (size_t)(i - 1) + (size_t)ix * (j - 1), p10 = p00 + 1, p01 = p00 + ix, p11 = p01 + 1;
    const double *c = a.cell;
    const double dxim = a.mx.dim[i - 1], dyim = a.my.dim[j - 1];
    double un[6];
#pragma unroll
    for (int q = 0; q < 6; q++) {   // mlwhalf.f:45-52
        const double *u = c + (size_t)q * N;
        const double uh = 0.25 * (u[p10] + u[p00] + u[p11] + u[p01]);
        const int sx = kFx[q], sy = kFy[q];
        const double dfdx = dxim * (fl(c, N, sx, p10) - fl(c, N, sx, p00) + fl(c, N, sx, p11) - fl(c, N, sx, p01)) / 2.0;
        const double dfdy = dyim * (fl(c, N, sy, p01) - fl(c, N, sy, p00) + fl(c, N, sy, p11) - fl(c, N, sy, p10)) / 2.0;
        un[q] = uh - dt * dfdx - dt * dfdy;
        if (q == 2 || q == 3) {   // mlwsrch.f: corner term of the rx, ry sources
            const double *s = c + (size_t)(q == 2 ? S_SRX : S_SRY) * N;
            const double sh = (s[p10] + s[p00] + s[p11] + s[p01]) / 4.0;
            un[q] = un[q] + dt * sh;
        }
    }
    const double gm = a.gm;
    const double r = a.xm[i - 1], st = a.trig[3 * jx + j - 1], ct = a.trig[4 * jx + j - 1];
    const double r2 = r * r;
    const double roh = un[0] / r2 / st;                                 // :234-247
    const double eeh = un[1] / r2 / st;
    const double vxh = un[2] / roh / r2 / st;
    const double vyh = un[3] / roh / r / st;
    const double bxh = un[4] / r2 / st;
    const double byh = un[5] / r / st;
    const double vv = vxh * vxh + vyh * vyh;
    const double bb = bxh * bxh + byh * byh;
    const double prh = (gm - 1.0) * (eeh - 0.5 * roh * vv - pi8i * bb);
    const double ezh = -vxh * byh + vyh * bxh;
    double *h = a.corner;
#define W(slot, v) h[(size_t)(slot) * N + p00] = (v)
    W(K_XRO, roh * vxh * r2 * st);                                      // :258-265
    W(K_YRO, roh * vyh * r * st);
    const double eph = prh * gm / (gm - 1.0) + 0.5 * roh * vv;         // :268-278
    W(K_XEE, (eph * vxh - byh * ezh * pi4i) * r2 * st);
    W(K_YEE, (eph * vyh + bxh * ezh * pi4i) * r * st);
    const double rxy = roh * vxh * vyh - pi4i * bxh * byh;              // :282-293
    W(K_XRX, (roh * (vxh * vxh) + prh + pi8i * (byh * byh - bxh * bxh)) * r2 * st);
    W(K_YRX, rxy * r * st);
    W(K_SRX, (roh * (vyh * vyh) + 2.0 * prh + pi4i * (bxh * bxh)) * r * st);
    W(K_XRY, (roh * vyh * vxh - pi4i * byh * bxh) * st);               // :298-310 (no radius, as the reference)
    W(K_YRY, (roh * (vyh * vyh) + prh + pi8i * (bxh * bxh - byh * byh)) * st);
    W(K_SRY, rxy * (-2.0) * st + (prh + pi8i * (bxh * bxh + byh * byh)) * ct);
    W(K_YBX, ezh * r * st);                                             // :315-321
    W(K_XBY, -ezh * r * st);                                            // :325-331
    W(K_SAZ, -ezh);                                                     // :335-340
#undef W
}

// interior cell (i, j) in [2, n-1]^2
__global__ __launch_bounds__(256) void mlw2s_update(Mlw2sArgs a) {
    double dt;
    if (!step_dt2s(a, dt)) return;
    const int ix = a.ix, jx = a.jx, ni = ix - 2;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)ni * (jx - 2)) return;
    const int i = 2 + (int)(t % ni), j = 2 + (int)(t / ni);
    const size_t N = (size_t)ix * jx;
    const size_t p = (size_t)(i - 1) + (size_t)ix * (j - 1);
    const size_t sy = ix;
    const double *c = a.cell, *h = a.corner;
    const double dxi = a.mx.di[i - 1], dyi = a.my.di[j - 1];
    const double dxim = a.mx.dim[i - 1], dyim = a.my.dim[j - 1];
    const double dxim1 = a.mx.dim[i - 2], dyim1 = a.my.dim[j - 2];
    const double ux0 = a.mx.u0[i - 1], ux1 = a.mx.u1[i - 1];
    const double uy0 = a.my.u0[j - 1], uy1 = a.my.u1[j - 1];
    const size_t c00 = p - 1 - sy, c10 = c00 + 1, c01 = c00 + sy, c11 = c01 + 1;   // corners (i-1|i, j-1|j)
    const double zero = 0.0;                                            // :346-357
    const double *vxp = a.prim[2], *vyp = a.prim[3];
    const double qxc = a.qav * a.dxm[i - 1] * fmaxF(zero, __builtin_fabs(vxp[p + 1] - vxp[p]) - kR4_1em4);
    const double qxw = a.qav * a.dxm[i - 2] * fmaxF(zero, __builtin_fabs(vxp[p] - vxp[p - 1]) - kR4_1em4);
    const double qyc = a.qav * a.dym[j - 1] * fmaxF(zero, __builtin_fabs(vyp[p + sy] - vyp[p]) - kR4_1em4);
    const double qys = a.qav * a.dym[j - 2] * fmaxF(zero, __builtin_fabs(vyp[p] - vyp[p - sy]) - kR4_1em4);
    double U[6];
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const int fx = kFx[q], fy = kFy[q], hx = kHx[q], hy = kHy[q];
        double du = 0.0;   // mlwhalf.f:36-41
        du = du - 0.5 * dt * (0.5 * dxi * (fl(c, N, fx, p + 1) - fl(c, N, fx, p - 1)))
                - 0.5 * dt * (0.5 * dyi * (fl(c, N, fy, p + sy) - fl(c, N, fy, p - sy)));
        if (q == 2 || q == 3) du = du + 0.5 * dt * c[(size_t)(q == 2 ? S_SRX : S_SRY) * N + p];   // mlwsrch.f
        const double dfdx = dxi * (uy0 * (fl(h, N, hx, c10) - fl(h, N, hx, c00)) + uy1 * (fl(h, N, hx, c11) - fl(h, N, hx, c01)));
        const double dfdy = dyi * (ux0 * (fl(h, N, hy, c01) - fl(h, N, hy, c00)) + ux1 * (fl(h, N, hy, c11) - fl(h, N, hy, c10)));
        du = du - 0.5 * dt * dfdx - 0.5 * dt * dfdy;                    // mlwfull.f
        if (q == 2 || q == 3) {                                         // mlwsrcf.f
            const double *s = h + (size_t)(q == 2 ? K_SRX : K_SRY) * N;
            const double sh = uy0 * (ux0 * s[c00] + ux1 * s[c10]) + uy1 * (ux0 * s[c01] + ux1 * s[c11]);
            du = du + 0.5 * dt * sh;
        }
        const double *u = c + (size_t)q * N;                            // mlwartv.f on the scaled variables
        const double uc = u[p];
        du = du + dt * (dxi * (qxc * dxim * (u[p + 1] - uc) - qxw * dxim1 * (uc - u[p - 1])))
                + dt * (dyi * (qyc * dyim * (u[p + sy] - uc) - qys * dyim1 * (uc - u[p - sy])));
        U[q] = uc + du;
    }
    double daz = 0.0;                                                   // Az: mlwsrch then mlwsrcf, s = -ez
    daz = daz + 0.5 * dt * c[(size_t)S_SAZ * N + p];
    {
        const double *s = h + (size_t)K_SAZ * N;
        const double sh = uy0 * (ux0 * s[c00] + ux1 * s[c10]) + uy1 * (ux0 * s[c01] + ux1 * s[c11]);
        daz = daz + 0.5 * dt * sh;
    }
    const double gm = a.gm;
    const double r = a.x[i - 1], st = a.trig[j - 1], r2 = r * r;        // :369-396
    const double ro = U[0] / r2 / st;
    const double ee = U[1] / r2 / st;
    const double vx = U[2] / ro / r2 / st;
    const double vy = U[3] / ro / r / st;
    const double bx = U[4] / r2 / st;
    const double by = U[5] / r / st;
    const double vv = vx * vx + vy * vy;
    const double bb = bx * bx + by * by;
    a.out[0][p] = ro;
    a.out[1][p] = (gm - 1.0) * (ee - 0.5 * ro * vv - pi8i * bb);
    a.out[2][p] = vx;
    a.out[3][p] = vy;
    a.out[4][p] = bx;
    a.out[5][p] = by;
    a.out[6][p] = a.prim[6][p] + daz;
}

}  // namespace

hipError_t launch_mlw2d_s(const Mlw2sArgs &a, hipStream_t s) {
    if (a.ix < 3 || a.jx < 3) return hipSuccess;
    if (!a.x || !a.xm || !a.trig) return hipErrorInvalidValue;
    const size_t N = (size_t)a.ix * a.jx;
    const size_t nc = (size_t)(a.ix - 1) * (a.jx - 1), nu = (size_t)(a.ix - 2) * (a.jx - 2);
    hipLaunchKernelGGL(mlw2s_cells, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(mlw2s_corners, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(mlw2s_update, dim3((unsigned)((nu + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace cans
