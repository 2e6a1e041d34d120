// cans_batch1d.hip -- many independent 1D MHD tubes, each running the whole
// CANS 1D main loop (cans1d/md_mhdshktb/hdmlw/main.f:132-186):
//     ns = ns+1; cfl_m; stop on merr; t = t+dt; mlw_m; bdfrex both ends;
//     loop while ns < nstop and t < tend
// with its own dt, on one MI355X.
//
// Persistent kernel: every workgroup takes tubes from a queue (the host
// orders it longest first) and runs a tube's loop to the end before taking
// the next, so tubes of very different lengths (nx 256..4096, dt differing
// by 10x) keep all CUs busy without a launch per step.  Inside a tube, a
// wave updates segments of 62 cells with the neighbours exchanged by DPP lane
// shifts; the free-boundary ghost fill (bdfrex.f:30-40) and the cfl_m term of
// the new state (cans1d/common/cfl_m.f:2-74) are fused into the update, so a
// step costs one workgroup barrier (the dt min-reduction).
//
// Precision: T = double is bit-identical to mlw_m.f:2-215 + cfl_m + bdfrex;
// T = float evaluates the cell and face fluxes (and their differences) in
// FP32 and keeps in FP64 the half-step state with its pressure recovery, the
// accumulation of du, the update, the primitive recovery and cfl_m.
#include "cans_common.hpp"
#include "cans_kernels.hpp"
#include "cans_wave.hpp"

namespace cans {
namespace {

constexpr double pi8i = 1.0 / kPi / 8.0;   // mlw_m.f:55
constexpr double pi4i = 1.0 / kPi / 4.0;   // mlw_m.f:56

// One wave updates the 62 interior cells c0..c0+61 of a tube (lane l holds
// cell c0-1+l): cur -> nxt, ghost copies, and the min of the new dtq terms.
struct P5 {
    double *q[5];
};

template <class T>
CANS_DEV double update_segment(const Batch1dArgs &a, const P5 &cur, const P5 &nxt, long long o, int ix, int c0,
                               double dt) {
    const int lane = threadIdx.x & 63;
    const int margin = a.margin;
    const int c = c0 - 1 + lane;                        // this lane's cell, 1-based
    const int cc = c < 1 ? 1 : (c > ix ? ix : c);       // clamped for the load
    const long long p = o + cc - 1;
    const double gm = a.gm;
    const double ro = cur.q[0][p], pr = cur.q[1][p], vx = cur.q[2][p], vy = cur.q[3][p], by = cur.q[4][p];
    const double bx = a.bx[p];
    // cells (mlw_m.f:79-125): conserved variables in FP64, fluxes in T
    double u[5];
    T f[5];
    {
        const double vv = vx * vx + vy * vy;
        const double bb = by * by;
        u[0] = ro;
        u[1] = pr / (gm - 1.0) + 0.5 * ro * vv + pi8i * bb;
        u[2] = ro * vx;
        u[3] = ro * vy;
        u[4] = by;
        const T tro = (T)ro, tpr = (T)pr, tvx = (T)vx, tvy = (T)vy, tby = (T)by, tbx = (T)bx, tgm = (T)gm;
        const T tvv = tvx * tvx + tvy * tvy;
        const T tbb = tby * tby;
        const T ez = -tvx * tby + tvy * tbx;
        const T ep = tpr * tgm / (tgm - (T)1.0) + (T)0.5 * tro * tvv;
        f[0] = tro * tvx;
        f[1] = ep * tvx - tby * ez * (T)pi4i;
        f[2] = tro * (tvx * tvx) + tpr + tbb * (T)pi8i;
        f[3] = tro * tvx * tvy - tbx * tby * (T)pi4i;
        f[4] = -ez;
    }
    const double dxi = a.di[p];
    const double dxim = a.dim[p];
    // face between this cell and the next (mlwhalf.f:39-46, mlw_m.f:129-175):
    // the half-step state and its primitives, pressure included, in FP64
    // (in the mixed form the flux difference that enters it is the FP32 one,
    // widened); the face fluxes in T.
    T ff[5];
    T fE[5];
    double uE[5];
    {
        double un[5];
#pragma unroll
        for (int q = 0; q < 5; q++) {
            uE[q] = shl1(u[q]);
            fE[q] = shl1(f[q]);
            const double uh = 0.5 * (uE[q] + u[q]);
            const double dfdx = dxim * (double)(fE[q] - f[q]);
            un[q] = uh - dt * dfdx;
        }
        const double roh = un[0], eeh = un[1], rxh = un[2], ryh = un[3], byh = un[4];
        const double bxm = a.bxm[p];
        const double vxh = rxh / roh;
        const double vyh = ryh / roh;
        const double vv = vxh * vxh + vyh * vyh;
        const double bb = byh * byh;
        const double prh = (gm - 1.0) * (eeh - 0.5 * roh * vv - pi8i * bb);
        const T troh = (T)roh, tvxh = (T)vxh, tvyh = (T)vyh, tbyh = (T)byh, tbxm = (T)bxm, tprh = (T)prh,
                tgm = (T)gm;
        const T tvv = tvxh * tvxh + tvyh * tvyh;
        const T tbb = tbyh * tbyh;
        const T ezh = -tvxh * tbyh + tvyh * tbxm;
        const T eph = tprh * tgm / (tgm - (T)1.0) + (T)0.5 * troh * tvv;
        ff[0] = troh * tvxh;
        ff[1] = eph * tvxh - tbyh * ezh * (T)pi4i;
        ff[2] = troh * (tvxh * tvxh) + tprh + tbb * (T)pi8i;
        ff[3] = troh * tvxh * tvyh - tbxm * tbyh * (T)pi4i;
        ff[4] = -ezh;
    }
    // own cell: centre term, full step, artificial viscosity, update
    const double zero = 0.0;
    const double vxE = shl1(vx);
    const double qxc = a.qav * a.dxm[p] * fmaxF(zero, __builtin_fabs(vxE - vx) - kR4_1em4);
    const double qxw = shr1(qxc);            // qav*dxm(c-1)*max(0,|vx(c)-vx(c-1)|-1e-4)
    const double dxim1 = shr1(dxim);
    double du[5];
#pragma unroll
    for (int q = 0; q < 5; q++) {
        du[q] = 0.0;
        const T fW = shr1(f[q]);
        du[q] = du[q] - 0.5 * dt * (0.5 * dxi * (double)(fE[q] - fW));      // mlwhalf.f:33-35
        const T dfdx = (T)dxi * (ff[q] - shr1(ff[q]));                      // mlwfull.f:30-33
        du[q] = du[q] - 0.5 * dt * (double)dfdx;
        const double dEc = uE[q] - u[q];                                    // mlwartv.f:31-34
        const double dcW = shr1(dEc);
        du[q] = du[q] + dt * (dxi * (qxc * dxim * dEc - qxw * dxim1 * dcW));
    }
    const double nro = u[0] + du[0];
    const double nrx = u[2] + du[2];
    const double nry = u[3] + du[3];
    const double nby = u[4] + du[4];
    const double nee = u[1] + du[1];
    const double nvx = nrx / nro;
    const double nvy = nry / nro;
    const double nvv = nvx * nvx + nvy * nvy;
    const double nbb = nby * nby;
    const double npr = (gm - 1.0) * (nee - 0.5 * nro * nvv - pi8i * nbb);

    double m = kR4_1e20;
    const bool out = lane >= 1 && lane <= 62 && c <= ix - margin;
    if (out) {
        nxt.q[0][p] = nro; nxt.q[1][p] = npr; nxt.q[2][p] = nvx; nxt.q[3][p] = nvy; nxt.q[4][p] = nby;
        // cfl_m runs over i = 2..ix-1 of the new state
        if (c >= 2 && c <= ix - 1) {
            const double q = dtq1(nro, npr, nvx, nby, bx, a.dx[p], gm, a.safety);
            if (q < m) m = q;
        }
        // bdfrex on both sides: ghosts take the first / last interior cell
        // (bdfrex.f:30-40), and cfl_m sees the ghosts 2..margin, ix-margin+1..ix-1
        for (int side = 0; side < 2; side++) {
            if (c != (side == 0 ? margin + 1 : ix - margin)) continue;
            for (int g = 1; g <= margin; g++) {
                const int gc = side == 0 ? c - g : c + g;
                const long long pg = o + gc - 1;
                nxt.q[0][pg] = nro; nxt.q[1][pg] = npr; nxt.q[2][pg] = nvx; nxt.q[3][pg] = nvy; nxt.q[4][pg] = nby;
                if (gc >= 2 && gc <= ix - 1) {
                    const double q = dtq1(nro, npr, nvx, nby, a.bx[pg], a.dx[pg], gm, a.safety);
                    if (q < m) m = q;
                }
            }
        }
    }
    return m;
}

template <class T, int BT>
__global__ __launch_bounds__(BT) void batch1d_kernel(Batch1dArgs a) {
    constexpr int NW = BT / 64;
    __shared__ double red[2][NW];
    __shared__ int tube_s;
    const int wv = uni((int)(threadIdx.x >> 6));
    for (;;) {
        // Queue pop by wave 0 under a uniform branch: lane 0 adds 1, the
        // other lanes add 0, every lane stores lane 0's ticket.
        if (wv == 0) {
            const unsigned got = atomicAdd(a.next, threadIdx.x == 0 ? 1u : 0u);
            tube_s = uni((int)got);
        }
        __syncthreads();
        const int kq = uni(tube_s);
        __syncthreads();
        if (kq >= a.ntube) break;
        const int k = a.order ? uni(a.order[kq]) : kq;
        const long long o = uni(a.off[k]);
        const int ix = uni(a.ix[k]);
        const int margin = a.margin;
        const P5 A{{a.f[0], a.f[1], a.f[2], a.f[3], a.f[4]}};
        const P5 B{{a.g[0], a.g[1], a.g[2], a.g[3], a.g[4]}};
        // cfl_m of the initial state over i = 2..ix-1
        double m = kR4_1e20;
        for (int c = 2 + (int)threadIdx.x; c <= ix - 1; c += BT) {
            const long long p = o + c - 1;
            const double q = dtq1(A.q[0][p], A.q[1][p], A.q[2][p], A.q[4][p], a.bx[p], a.dx[p], a.gm, a.safety);
            if (q < m) m = q;
        }
        int par = 0;
        double dt = block_min<BT>(m, red[par]);
        par ^= 1;
        int ns = 0, merr = 0, nexec = 0;
        double t = a.t0;
        const int nseg = (ix - 2 * margin + 61) / 62;
        const int nstop = a.nstop_tube ? uni(a.nstop_tube[k]) : a.nstop;
        for (;;) {
            ns = ns + 1;
            if (dt < a.dtmin) { merr = 9001; break; }    // cfl_m.f: merr=9001
            t = t + dt;
            const P5 cur = (nexec & 1) ? B : A;
            const P5 nxt = (nexec & 1) ? A : B;
            m = kR4_1e20;
            for (int s = wv; s < nseg; s += NW) {
                const double q = update_segment<T>(a, cur, nxt, o, ix, margin + 1 + 62 * s, dt);
                if (q < m) m = q;
            }
            nexec++;
            dt = block_min<BT>(m, red[par]);              // dt of the next step; also the step's barrier
            par ^= 1;
            if (!(ns < nstop && t < a.tend)) break;
        }
        if (nexec & 1) {   // the last state is in the scratch set: copy back
            for (int c = 1 + (int)threadIdx.x; c <= ix; c += BT) {
                const long long p = o + c - 1;
#pragma unroll
                for (int q = 0; q < 5; q++) A.q[q][p] = B.q[q][p];
            }
        }
        if (wv == 0) {   // uniform branch; the lanes store identical values
            a.t_out[k] = t;
            a.ns_out[k] = ns;
            a.merr_out[k] = merr;
        }
    }
}

__global__ void batch1d_prep(const double *dx, const double *dxm, double *di, double *dim, long long n,
                             unsigned *next) {
    const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i0 == 0) *next = 0u;
    for (long long i = i0; i < n; i += (long long)gridDim.x * blockDim.x) {
        di[i] = 1.0 / dx[i];
        dim[i] = 1.0 / dxm[i];
    }
}

template <class T>
hipError_t launch_bt(const Batch1dArgs &a, int bt, int nblocks, hipStream_t s) {
    switch (bt) {
    case 64: hipLaunchKernelGGL((batch1d_kernel<T, 64>), dim3(nblocks), dim3(64), 0, s, a); break;
    case 128: hipLaunchKernelGGL((batch1d_kernel<T, 128>), dim3(nblocks), dim3(128), 0, s, a); break;
    case 256: hipLaunchKernelGGL((batch1d_kernel<T, 256>), dim3(nblocks), dim3(256), 0, s, a); break;
    case 512: hipLaunchKernelGGL((batch1d_kernel<T, 512>), dim3(nblocks), dim3(512), 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_batch1d(const Batch1dArgs &a, int mixed, int block_threads, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(batch1d_prep, dim3(1024), dim3(256), 0, s, a.dx, a.dxm, a.di, a.dim, a.ncells, a.next);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return mixed ? launch_bt<float>(a, block_threads, nblocks, s) : launch_bt<double>(a, block_threads, nblocks, s);
}

}  // namespace cans
