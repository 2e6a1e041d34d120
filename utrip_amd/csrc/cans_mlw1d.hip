// cans_mlw1d.hip -- fused 1D modified Lax-Wendroff MHD step (gfx950).
//
// One launch does cans1d/hdmlw/mlw_m.f:2-215 (mlwhalf, mlwfull, mlwartv,
// update, primitive recovery).  A block of NT threads owns NT-2 interior
// cells: every thread forms the conserved variables and cell fluxes of one
// loaded cell, then the half-step state of the face to its right, both
// through LDS, and finally the update of its own cell.  Bx enters only the
// induction/momentum fluxes, not the energy (mlw_m.f:79-86).
#include "cans_common.hpp"
#include "cans_clock.hpp"
#include "cans_kernels.hpp"

namespace cans {
namespace {

constexpr int NT = 256;
constexpr int NI = NT - 2;                 // interior cells per block
constexpr double pi8i = 1.0 / kPi / 8.0;   // mlw_m.f:55
constexpr double pi4i = 1.0 / kPi / 4.0;   // mlw_m.f:56

__global__ __launch_bounds__(NT) void mlw1d_kernel(Mlw1dArgs a) {
    __shared__ double u[5][NT];    // ro ee rx ry by at the loaded cells
    __shared__ double f[5][NT];    // cell fluxes, then face fluxes
    __shared__ double vxs[NT];

    const int ix = a.ix;
    const int tid = threadIdx.x;
    const int ib = 2 + blockIdx.x * NI;      // first interior cell of the block
    const int ic = ib - 1 + tid;             // the cell this thread loads (1-based)

    double dt = a.dt;
    if (a.clk) {
        StepCtl c = clock_use(a.clk, a.step, a.peek != 0, blockIdx.x == 0 && tid == 0);
        if (!c.run) return;
        dt = c.dt;
    }
    const double gm = a.gm;

    // cells: conserved variables and fluxes (mlw_m.f:79-125)
    if (ic <= ix) {
        const int p = ic - 1;
        const double ro = a.ro[p], pr = a.pr[p], vx = a.vx[p], vy = a.vy[p], by = a.by[p];
        const double bx = a.bx[p];
        const double vv = vx * vx + vy * vy;
        const double bb = by * by;
        const double ez = -vx * by + vy * bx;
        const double ep = pr * gm / (gm - 1.0) + 0.5 * ro * vv;
        u[0][tid] = ro;
        u[1][tid] = pr / (gm - 1.0) + 0.5 * ro * vv + pi8i * bb;
        u[2][tid] = ro * vx;
        u[3][tid] = ro * vy;
        u[4][tid] = by;
        vxs[tid] = vx;
        f[0][tid] = ro * vx;
        f[1][tid] = ep * vx - by * ez * pi4i;
        f[2][tid] = ro * (vx * vx) + pr + bb * pi8i;
        f[3][tid] = ro * vx * vy - bx * by * pi4i;
        f[4][tid] = -ez;
    }
    __syncthreads();

    // own interior cell
    const int i = ic;                        // owned when 1 <= tid <= NT-2
    const bool own = tid >= 1 && tid <= NT - 2 && i <= ix - 1;
    double du[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double dxi = 0.0;
    if (own) {
        dxi = a.mx.di[i - 1];
#pragma unroll
        for (int q = 0; q < 5; q++)   // mlwhalf.f:33-35
            du[q] = du[q] - 0.5 * dt * (0.5 * dxi * (f[q][tid + 1] - f[q][tid - 1]));
    }

    // face between this cell and the next (mlwhalf.f:39-46, mlw_m.f:129-175)
    double ff[5];
    const bool face = tid <= NT - 2 && ic <= ix - 1;
    if (face) {
        const double dxim = a.mx.dim[ic - 1];
        double un[5];
#pragma unroll
        for (int q = 0; q < 5; q++) {
            const double uh = 0.5 * (u[q][tid + 1] + u[q][tid]);
            const double dfdx = dxim * (f[q][tid + 1] - f[q][tid]);
            un[q] = uh - dt * dfdx;
        }
        const double roh = un[0], eeh = un[1], rxh = un[2], ryh = un[3], byh = un[4];
        const double bxm = a.bxm[ic - 1];
        const double vxh = rxh / roh;
        const double vyh = ryh / roh;
        const double vv = vxh * vxh + vyh * vyh;
        const double bb = byh * byh;
        const double prh = (gm - 1.0) * (eeh - 0.5 * roh * vv - pi8i * bb);
        const double ezh = -vxh * byh + vyh * bxm;
        const double eph = prh * gm / (gm - 1.0) + 0.5 * roh * vv;
        ff[0] = roh * vxh;
        ff[1] = eph * vxh - byh * ezh * pi4i;
        ff[2] = roh * (vxh * vxh) + prh + bb * pi8i;
        ff[3] = roh * vxh * vyh - bxm * byh * pi4i;
        ff[4] = -ezh;
    }
    __syncthreads();
    if (face) {
#pragma unroll
        for (int q = 0; q < 5; q++) f[q][tid] = ff[q];
    }
    __syncthreads();
    if (!own) return;

    // full step (mlwfull.f:30-33)
#pragma unroll
    for (int q = 0; q < 5; q++) {
        const double dfdx = dxi * (f[q][tid] - f[q][tid - 1]);
        du[q] = du[q] - 0.5 * dt * dfdx;
    }
    // artificial viscosity (mlw_m.f:180-191, mlwartv.f:31-34)
    const double zero = 0.0;
    const double dxim = a.mx.dim[i - 1], dxim1 = a.mx.dim[i - 2];
    const double qxc = a.qav * a.dxm[i - 1] * fmaxF(zero, __builtin_fabs(vxs[tid + 1] - vxs[tid]) - kR4_1em4);
    const double qxw = a.qav * a.dxm[i - 2] * fmaxF(zero, __builtin_fabs(vxs[tid] - vxs[tid - 1]) - kR4_1em4);
#pragma unroll
    for (int q = 0; q < 5; q++)
        du[q] = du[q] + dt * (dxi * (qxc * dxim * (u[q][tid + 1] - u[q][tid])
                                     - qxw * dxim1 * (u[q][tid] - u[q][tid - 1])));
    // update and primitives (mlw_m.f:195-212)
    const int p = i - 1;
    const double ro = u[0][tid] + du[0];
    const double rx = u[2][tid] + du[2];
    const double ry = u[3][tid] + du[3];
    const double by = u[4][tid] + du[4];
    const double ee = u[1][tid] + du[1];
    const double vx = rx / ro;
    const double vy = ry / ro;
    const double vv = vx * vx + vy * vy;
    const double bb = by * by;
    a.ro_o[p] = ro;
    a.vx_o[p] = vx;
    a.vy_o[p] = vy;
    a.by_o[p] = by;
    a.pr_o[p] = (gm - 1.0) * (ee - 0.5 * ro * vv - pi8i * bb);
}

}  // namespace

hipError_t launch_mlw1d(const Mlw1dArgs &a, hipStream_t s) {
    if (a.ix < 3) return hipSuccess;
    const int nb = (a.ix - 2 + NI - 1) / NI;
    hipLaunchKernelGGL(mlw1d_kernel, dim3(nb), dim3(NT), 0, s, a);
    return hipGetLastError();
}

}  // namespace cans
