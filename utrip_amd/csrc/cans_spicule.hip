// cans_spicule.hip -- the UTRIP spicule model on gfx950.
//
//  * mlw1d_bg_kernel: one call of mlw_m_bg (cans1d/hdmlw/mlw_m_bg.f:2-271),
//    the modified Lax-Wendroff step of 1D MHD in a flux tube of
//    cross-section sc bent at radius rr, with gravity gx.
//  * spicule_kernel: the whole time loop of md_spicule/hdmlw/main.f:169-227
//      ns=ns+1; cfl_m; t=t+dt; mlw_m_bg; trqalfven; bnd; chkdav(ro); chkdav(pr)
//    for any number of independent realizations, one workgroup each, in one
//    persistent kernel (a queue hands realizations to workgroups).
//
// Inside a realization a wave updates a segment of 62 cells: lane l holds
// cell c0-1+l, the face between a cell and its east neighbour is computed by
// the cell's lane, neighbours travel by DPP shifts.  In the loop kernel the
// torque, the ghost images of bnd.f, both chkdav calls and the cfl_m term of
// the new state are fused into the update, so a step is one workgroup
// barrier (the dt min).  All arithmetic is FP64 in the reference's order:
// bit-identical to the Fortran compiled without FMA contraction.
//
// The torque's noise (RANDOM_NUMBER in trqalfven.f:24) is the Park-Miller
// stream of common/rangen.f: call n of trqalfven draws one number per cell
// in cell order, so cell i uses seed_n * 16807^i mod (2^31-1), and the seed
// advances by 16807^ix per call.  Lanes compute their draw directly from the
// call's starting seed with a 64-bit product folded mod 2^31-1.
#include "cans_common.hpp"
#include "cans_kernels.hpp"
#include "cans_wave.hpp"

namespace cans {
namespace {

constexpr double pi8i = 1.0 / kPi / 8.0;   // mlw_m_bg.f:70
constexpr double pi4i = 1.0 / kPi / 4.0;   // mlw_m_bg.f:71
constexpr unsigned long long kPmM = 2147483647ull;
constexpr long long kPmA = 16807;
constexpr double kPmRnml = 4.656612875245796924105750827168e-10;   // rangen.f:15
constexpr double kTrq0 = 30.0;                                      // trqalfven.f:33

// (s * p) mod (2^31-1) for s, p < 2^31: since 2^31 = 1 mod (2^31-1), two
// folds of the 62-bit product leave a value in [0, 2^31-1].
CANS_DEV long long pm_mul(long long s, long long p) {
    unsigned long long x = (unsigned long long)s * (unsigned long long)p;
    x = (x & kPmM) + (x >> 31);
    x = (x & kPmM) + (x >> 31);
    return x == kPmM ? 0 : (long long)x;
}

struct SpGeo {
    const double *g[SG_N];
    const double *di, *dim, *ux0, *ux1;
};

CANS_DEV SpGeo geo_at(const double *const *geo, const double *const *met, long long off) {
    SpGeo G;
#pragma unroll
    for (int k = 0; k < SG_N; k++) G.g[k] = geo[k] + off;
    G.di = met[0] + off;
    G.dim = met[1] + off;
    G.ux0 = met[2] + off;
    G.ux1 = met[3] + off;
    return G;
}

struct BgNew {
    double ro, pr, vx, vy, by;
};

// mlw_m_bg for the cell c = c0-1+lane of this lane; the result is the
// reference's for every lane 1..62 whose cell lies in 2..ix-1.  st: the
// state (5 fields) at element offset so.
CANS_DEV BgNew bg_update(const double *const *st, long long so, const SpGeo &G, int ix, int c, double dt, double qav,
                         double gm) {
    const int cc = c < 1 ? 1 : (c > ix ? ix : c);
    const int p = cc - 1;
    const long long ps = so + p;
    const double ro = st[0][ps], pr = st[1][ps], vx = st[2][ps], vy = st[3][ps], by = st[4][ps];
    const double bx = G.g[SG_BX][p], sc = G.g[SG_SC][p], rr = G.g[SG_RR][p];
    // cell values, fluxes and sources (mlw_m_bg.f:97-167)
    double u[5], f[5], s1, s2;
    {
        const double vv = vx * vx + vy * vy;
        const double bb = by * by + bx * bx;
        const double ee = pr / (gm - 1.0) + 0.5 * ro * vv + pi8i * bb;
        const double ez = -vx * by + vy * bx;
        const double gx = G.g[SG_GX][p];
        u[0] = ro * sc;
        u[1] = ee * sc;
        u[2] = (ro * vx) * sc;
        u[3] = (ro * vy) * sc * rr;
        u[4] = by * sc / rr;
        f[0] = (ro * vx) * sc;
        const double ep = pr * gm / (gm - 1.0) + 0.5 * ro * vv;
        f[1] = (ep * vx - by * ez * pi4i) * sc;
        s1 = ro * vx * gx * sc;
        const double bbm = by * by - bx * bx;
        f[2] = (ro * (vx * vx) + pr + bbm * pi8i) * sc;
        s2 = (ro * gx + (ro * (vy * vy) - (by * by) * pi4i) / rr * G.g[SG_DRR][p]) * sc + (pr + bb * pi8i) * G.g[SG_DSC][p];
        f[3] = (ro * vx * vy - bx * by * pi4i) * sc * rr;
        f[4] = -ez * sc / rr;
    }
    // the face between this cell and the next: half step (mlwhalf.f:38-42,
    // mlwsrch.f:38-41), half-step primitives and fluxes (mlw_m_bg.f:170-233)
    const double dxim = G.dim[p];
    double uE[5], fE[5], ff[5], sf1, sf2;
    {
        double un[5];
#pragma unroll
        for (int q = 0; q < 5; q++) {
            uE[q] = shl1(u[q]);
            fE[q] = shl1(f[q]);
            const double uh = 0.5 * (uE[q] + u[q]);
            const double dfdx = dxim * (fE[q] - f[q]);
            un[q] = uh - dt * dfdx;
        }
        un[1] = un[1] + dt * (0.5 * (shl1(s1) + s1));
        un[2] = un[2] + dt * (0.5 * (shl1(s2) + s2));
        const double scm = G.g[SG_SCM][p], rrm = G.g[SG_RRM][p], bxm = G.g[SG_BXM][p], gxm = G.g[SG_GXM][p];
        const double roh = un[0] / scm;
        const double eeh = un[1] / scm;
        const double vxh = un[2] / un[0];
        const double vyh = un[3] / un[0] / rrm;
        const double byh = un[4] / scm * rrm;
        const double vv = vxh * vxh + vyh * vyh;
        const double bbh = bxm * bxm + byh * byh;
        const double prh = (gm - 1.0) * (eeh - 0.5 * roh * vv - pi8i * bbh);
        const double ezh = -vxh * byh + vyh * bxm;
        ff[0] = (roh * vxh) * scm;
        const double eph = prh * gm / (gm - 1.0) + 0.5 * roh * vv;
        ff[1] = (eph * vxh - byh * ezh * pi4i) * scm;
        sf1 = roh * vxh * gxm * scm;
        const double bbm = byh * byh - bxm * bxm;
        const double bb = byh * byh + bxm * bxm;
        ff[2] = (roh * (vxh * vxh) + prh + bbm * pi8i) * scm;
        sf2 = (roh * gxm + (roh * (vyh * vyh) - (byh * byh) * pi4i) / rrm * G.g[SG_DRRM][p]) * scm
            + (prh + bb * pi8i) * G.g[SG_DSCM][p];
        ff[3] = (roh * vxh * vyh - bxm * byh * pi4i) * scm * rrm;
        ff[4] = -ezh * scm / rrm;
    }
    // own cell: centre term, sources, full step, viscosity (mlwhalf.f:33-35,
    // mlwsrch.f:32-34, mlwfull.f:30-33, mlwsrcf.f:31-34, mlwartv.f:31-34)
    const double dxi = G.di[p];
    const double ux0 = G.ux0[p], ux1 = G.ux1[p];
    const double zero = 0.0;
    const double qxc = qav * G.g[SG_DXM][p] * fmaxF(zero, __builtin_fabs(shl1(vx) - vx) - kR4_1em4);
    const double qxw = shr1(qxc);
    const double dxim1 = shr1(dxim);
    const double sfW1 = shr1(sf1), sfW2 = shr1(sf2);
    double du[5];
#pragma unroll
    for (int q = 0; q < 5; q++) {
        du[q] = 0.0;
        du[q] = du[q] - 0.5 * dt * (0.5 * dxi * (fE[q] - shr1(f[q])));
        if (q == 1) du[q] = du[q] + 0.5 * dt * s1;
        if (q == 2) du[q] = du[q] + 0.5 * dt * s2;
        const double dfdx = dxi * (ff[q] - shr1(ff[q]));
        du[q] = du[q] - 0.5 * dt * dfdx;
        if (q == 1) du[q] = du[q] + 0.5 * dt * (ux0 * sfW1 + ux1 * sf1);
        if (q == 2) du[q] = du[q] + 0.5 * dt * (ux0 * sfW2 + ux1 * sf2);
        const double dEc = uE[q] - u[q];
        const double dcW = shr1(dEc);
        du[q] = du[q] + dt * (dxi * (qxc * dxim * dEc - qxw * dxim1 * dcW));
    }
    // update and primitive recovery (mlw_m_bg.f:251-268)
    BgNew o;
    o.ro = (u[0] + du[0]) / sc;
    const double ee = (u[1] + du[1]) / sc;
    const double rx = (u[2] + du[2]) / sc;
    const double ry = (u[3] + du[3]) / sc / rr;
    o.by = (u[4] + du[4]) / sc * rr;
    o.vx = rx / o.ro;
    o.vy = ry / o.ro;
    const double vv = o.vx * o.vx + o.vy * o.vy;
    const double bb = bx * bx + o.by * o.by;
    o.pr = (gm - 1.0) * (ee - 0.5 * o.ro * vv - pi8i * bb);
    return o;
}

__global__ __launch_bounds__(64) void mlw1d_bg_kernel(MlwBgArgs a) {
    const int lane = threadIdx.x;
    const int c = 2 + 62 * (int)blockIdx.x - 1 + lane;
    const SpGeo G = geo_at(a.geo, a.met, 0);
    const BgNew n = bg_update(a.in, 0, G, a.ix, c, a.dt, a.qav, a.gm);
    if (lane >= 1 && lane <= 62 && c <= a.ix - 1) {
        const int p = c - 1;
        a.out[0][p] = n.ro; a.out[1][p] = n.pr; a.out[2][p] = n.vx; a.out[3][p] = n.vy; a.out[4][p] = n.by;
    }
}

// chkdav(ro) then chkdav(pr) on one cell (chkdav.f:33-38), the store, and
// the cell's cfl_m term when it lies in 2..ix-1
CANS_DEV void put_cell(const SpiculeArgs &a, double *const *nxt, long long so, const SpGeo &G, int ci, double ro,
                       double pr, double vx, double vy, double by, double &m) {
    if (ro < a.floor) { ro = a.floor; vx = 0.0; }
    if (pr < a.floor) { pr = a.floor; vx = 0.0; }
    const long long ps = so + ci - 1;
    nxt[0][ps] = ro; nxt[1][ps] = pr; nxt[2][ps] = vx; nxt[3][ps] = vy; nxt[4][ps] = by;
    if (ci >= 2 && ci <= a.ix - 1) {
        const double q = dtq1(ro, pr, vx, by, G.g[SG_BX][ci - 1], G.g[SG_DX][ci - 1], a.gm, a.safety);
        if (q < m) m = q;
    }
}

// One step of one 62-cell segment of a realization: mlw_m_bg, trqalfven,
// bnd.f (mirror at the foot point, vx odd; free at the top), chkdav; returns
// the min of the new cfl_m terms of the cells it wrote.
CANS_DEV double spicule_segment(const SpiculeArgs &a, const double *const *cur, double *const *nxt, long long so,
                                const SpGeo &G, const double *trq, int c0, double dt, long long seed) {
    const int lane = threadIdx.x & 63;
    const int ix = a.ix, margin = a.margin;
    const int c = c0 - 1 + lane;
    const BgNew n = bg_update(cur, so, G, ix, c, dt, a.qav, a.gm);
    double m = kR4_1e20;
    if (lane >= 1 && lane <= 62 && c <= ix - margin) {
        const int p = c - 1;
        // trqalfven.f:23-37: q = 2 rr (anum - 0.5), vy += dt trq0 q trq
        const double anum = kPmRnml * (double)pm_mul(seed, a.pw[p]);
        const double q = 2.0 * G.g[SG_RR][p] * (anum - 0.5);
        const double vy = n.vy + dt * kTrq0 * q * trq[p];
        put_cell(a, nxt, so, G, c, n.ro, n.pr, n.vx, vy, n.by, m);
        if (c <= 2 * margin)        // bdsppx / bdspnx(0, margin): ghost 2*margin+1-c
            put_cell(a, nxt, so, G, 2 * margin + 1 - c, n.ro, n.pr, -n.vx, vy, n.by, m);
        if (c == ix - margin)       // bdfrex(1, margin)
            for (int g = 1; g <= margin; g++) put_cell(a, nxt, so, G, c + g, n.ro, n.pr, n.vx, vy, n.by, m);
    }
    return m;
}

template <int BT>
__global__ __launch_bounds__(BT) void spicule_kernel(SpiculeArgs a) {
    constexpr int NW = BT / 64;
    __shared__ double red[2][NW];
    __shared__ int run_s;
    const int wv = uni((int)(threadIdx.x >> 6));
    const int ix = a.ix, margin = a.margin;
    const int nseg = (ix - 2 * margin + 61) / 62;
    const long long adv = a.pw[ix - 1];           // 16807^ix: one trqalfven call
    for (;;) {
        // queue pop by wave 0 under a uniform branch (lane 0 adds 1)
        if (wv == 0) {
            const unsigned got = atomicAdd(a.next, threadIdx.x == 0 ? 1u : 0u);
            run_s = uni((int)got);
        }
        __syncthreads();
        const int r = uni(run_s);
        __syncthreads();
        if (r >= a.nrun) break;
        const long long so = (long long)r * ix;
        const long long go = (long long)r * a.gstride;
        const SpGeo G = geo_at(a.geo, a.met, go);
        const double *trq = a.trq + go;
        double *A[5] = {a.f[0], a.f[1], a.f[2], a.f[3], a.f[4]};
        double *B[5] = {a.g[0], a.g[1], a.g[2], a.g[3], a.g[4]};
        // cfl_m of the state the loop starts from (cfl_m.f:46-60)
        double m = kR4_1e20;
        for (int c = 2 + (int)threadIdx.x; c <= ix - 1; c += BT) {
            const long long ps = so + c - 1;
            const double q = dtq1(A[0][ps], A[1][ps], A[2][ps], A[4][ps], G.g[SG_BX][c - 1], G.g[SG_DX][c - 1], a.gm,
                                  a.safety);
            if (q < m) m = q;
        }
        int par = 0;
        double dt = block_min<BT>(m, red[par]);
        par ^= 1;
        int ns = uni(a.ns[r]);
        int merr = 0, nexec = 0, nsnap = 0;
        double t = uni(a.t[r]);
        long long seed = uni((long long)a.seed[r]);
        for (;;) {
            ns = ns + 1;
            if (dt < a.dtmin) { merr = 9001; break; }   // cfl_m.f:64-71
            const double tp = t;
            t = t + dt;
            double *const *cur = (nexec & 1) ? B : A;
            double *const *nxt = (nexec & 1) ? A : B;
            m = kR4_1e20;
            for (int sg = wv; sg < nseg; sg += NW) {
                const double q = spicule_segment(a, cur, nxt, so, G, trq, margin + 1 + 62 * sg, dt, seed);
                if (q < m) m = q;
            }
            seed = pm_mul(seed, adv);
            nexec++;
            dt = block_min<BT>(m, red[par]);            // the next dt; also the step's barrier
            par ^= 1;
            // output when int(t/dtout) changes (main.f:203-206): the whole new
            // state is in nxt after the barrier; the next step only reads it
            if (a.snap && (int)(tp / a.dtout) < (int)(t / a.dtout) && nsnap < a.nsnap_max) {
                const long long k = (long long)r * a.nsnap_max + nsnap;
                for (int i = (int)threadIdx.x; i < ix; i += BT) {
#pragma unroll
                    for (int q = 0; q < 5; q++) a.snap[(k * 5 + q) * ix + i] = nxt[q][so + i];
                }
                if (wv == 0) a.snap_t[k] = t;
                nsnap++;
            }
            if (!(ns < a.nstop && t < a.tend)) break;
        }
        if (nexec & 1) {   // the last state is in the scratch set: copy it back
            for (int i = (int)threadIdx.x; i < ix; i += BT) {
#pragma unroll
                for (int q = 0; q < 5; q++) A[q][so + i] = B[q][so + i];
            }
        }
        if (wv == 0) {     // uniform branch; the lanes store identical values
            a.t[r] = t;
            a.ns[r] = ns;
            a.merr[r] = merr;
            a.seed[r] = (int)seed;
            if (a.nsnap) a.nsnap[r] = nsnap;
        }
        __syncthreads();   // the copy-back completes before the next pop
    }
}

// mlw_m_bg.f:75-82 per geometry set (1/dxm, 1/dx, ux0, ux1), the draw
// multipliers 16807^(i+1) mod (2^31-1) by square-and-multiply, the counter
__global__ void spicule_prep(const double *dx, const double *dxm, double *di, double *dim, double *ux0, double *ux1,
                             int ix, int nset, long long gstride, long long *pw, unsigned *next) {
    const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i0 == 0 && next) *next = 0u;
    const long long n = (long long)nset * ix;
    for (long long k = i0; k < n; k += (long long)gridDim.x * blockDim.x) {
        const long long set = k / ix;
        const int i = (int)(k - set * ix);
        const long long o = set * gstride + i;
        di[o] = 1.0 / dx[o];
        dim[o] = 1.0 / dxm[o];
        if (i >= 1 && i <= ix - 2) {
            ux1[o] = 0.5 * dxm[o - 1] / dx[o];
            ux0[o] = 0.5 * dxm[o] / dx[o];
        } else {
            ux1[o] = 0.0;
            ux0[o] = 0.0;
        }
    }
    if (pw) {
        for (long long i = i0; i < ix; i += (long long)gridDim.x * blockDim.x) {
            long long e = i + 1, b = kPmA, r = 1;
            while (e > 0) {
                if (e & 1) r = pm_mul(r, b);
                b = pm_mul(b, b);
                e >>= 1;
            }
            pw[i] = r;
        }
    }
}

// trqalfven.f:23-37 for one call; cell ix takes cell ix-2's draw and torque
__global__ void trq1d_kernel(double *vy, double dt, const double *trq, const double *rr, const long long *pw,
                             long long seed, int ix) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= ix) return;
    const int iq = (i == ix - 1) ? ix - 3 : i;
    const double anum = kPmRnml * (double)pm_mul(seed, pw[iq]);
    const double q = 2.0 * rr[iq] * (anum - 0.5);
    vy[i] = vy[i] + dt * kTrq0 * q * trq[i];
}

// chkdav.f:31-39 in one workgroup; the count is stored once
__global__ __launch_bounds__(1024) void chkdav1d_kernel(double *da, double *vx, double floor_, int n, int *count) {
    __shared__ int part[16];
    int c = 0;
    for (int i = (int)threadIdx.x; i < n; i += 1024) {
        if (da[i] < floor_) {
            da[i] = floor_;
            vx[i] = 0.0;
            c++;
        }
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x < 64) {
        int s = threadIdx.x < 16 ? part[threadIdx.x] : 0;
        for (int off = 8; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if (threadIdx.x == 0) *count = s;
    }
}

template <int BT>
hipError_t launch_sp(const SpiculeArgs &a, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(spicule_kernel<BT>, dim3(nblocks), dim3(BT), 0, s, a);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_mlw1d_bg(const MlwBgArgs &a, hipStream_t s) {
    const int nseg = (a.ix - 2 + 61) / 62;
    hipLaunchKernelGGL(mlw1d_bg_kernel, dim3(nseg), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_spicule_prep(const double *dx, const double *dxm, double *const *met, int ix, int nset,
                               long long gstride, long long *pw, unsigned *next, hipStream_t s) {
    const long long n = (long long)nset * ix;
    long long blocks = (n + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(spicule_prep, dim3((unsigned)blocks), dim3(256), 0, s, dx, dxm, met[0], met[1], met[2], met[3],
                       ix, nset, gstride, pw, next);
    return hipGetLastError();
}

hipError_t launch_spicule(const SpiculeArgs &a, int block_threads, int nblocks, hipStream_t s) {
    switch (block_threads) {
    case 64: return launch_sp<64>(a, nblocks, s);
    case 128: return launch_sp<128>(a, nblocks, s);
    case 256: return launch_sp<256>(a, nblocks, s);
    case 512: return launch_sp<512>(a, nblocks, s);
    case 1024: return launch_sp<1024>(a, nblocks, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_trq1d(double *vy, double dt, const double *trq, const double *rr, const long long *pw,
                        long long seed, int ix, hipStream_t s) {
    hipLaunchKernelGGL(trq1d_kernel, dim3((ix + 255) / 256), dim3(256), 0, s, vy, dt, trq, rr, pw, seed, ix);
    return hipGetLastError();
}

hipError_t launch_chkdav1d(double *da, double *vx, double floor_, int n, int *count, hipStream_t s) {
    hipLaunchKernelGGL(chkdav1d_kernel, dim3(1), dim3(1024), 0, s, da, vx, floor_, n, count);
    return hipGetLastError();
}

}  // namespace cans
