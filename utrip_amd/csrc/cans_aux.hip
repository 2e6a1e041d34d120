// cans_aux.hip -- small kernels around the engines: metric coefficients,
// ghost-zone fills (the CANS bc/ routines), the CFL time-step reduction
// (common/cfl_m.f in 1D/2D/3D), interior copies and the device clock.
#include "cans_common.hpp"
#include "cans_clock.hpp"
#include "cans_kernels.hpp"
#include "cans_wave.hpp"

namespace cans {
namespace {

// ---------------------------------------------------------------- metric
// mlw_m.f preamble, e.g. cans2d/hdmlw/mlw_m.f:68-84:
//   dxi(i)=1.0/dx(i); dxim(i)=1.0/dxm(i);
//   ux1(i)=0.5*dxm(i-1)/dx(i); ux0(i)=0.5*dxm(i)/dx(i)   (i=2..n-1)
__global__ void metric_kernel(const double *d, const double *dm, int n, Metric m) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;  // 1-based
    if (i > n) return;
    m.di[i - 1] = 1.0 / d[i - 1];
    m.dim[i - 1] = 1.0 / dm[i - 1];
    if (i >= 2 && i <= n - 1) {
        m.u1[i - 1] = 0.5 * dm[i - 2] / d[i - 1];
        m.u0[i - 1] = 0.5 * dm[i - 1] / d[i - 1];
    } else {
        m.u1[i - 1] = 0.0;
        m.u0[i - 1] = 0.0;
    }
}

// ---------------------------------------------------------------- boundary
// One thread per (field, line).  A line runs along `dir` through the array;
// the fill is the 1D index map of cans1d/bc/*.f applied to that line, which
// is what the 2D/3D routines do for every transverse index (including the
// transverse ghosts, cans2d/bc/bdperx.f:31-44).
__global__ void bnd_kernel(BndArgs a) {
    const int n = a.dir == 0 ? a.ix : (a.dir == 1 ? a.jx : a.kx);
    const long nlines = (long)a.ix * a.jx * a.kx / n;
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int f = (int)(gid / nlines);
    const long l = gid - (long)f * nlines;
    if (f >= a.nf) return;
    const int kind = a.kind[f];
    if (kind < 0) return;
    // base offset and stride of the line
    long base, stride;
    if (a.dir == 0) { base = l * a.ix; stride = 1; }
    else if (a.dir == 1) { const long i = l % a.ix, k = l / a.ix; base = i + k * (long)a.ix * a.jx; stride = a.ix; }
    else { base = l; stride = (long)a.ix * a.jx; }
    double *q = a.q[f] + base;
#define Q(idx) q[(long)((idx) - 1) * stride]
    const int margin = a.margin;
    if (kind == BD_PER) {   // bdperx.f:31-40
        const int margin0 = a.mbnd, margin1 = a.margin;
        const int ibnd0 = 1 + margin0, ibnd1 = n - margin1;
        for (int i = 1; i <= margin0; i++) Q(ibnd0 - i) = Q(ibnd1 + 1 - i);
        for (int i = 1; i <= margin1; i++) Q(ibnd1 + i) = Q(ibnd0 - 1 + i);
        return;
    }
    if (a.mbnd == 0) {
        const int ibnd = 1 + margin;
        double dqq = 0.0;
        if (kind == BD_FRD) dqq = (Q(ibnd + 1) - Q(ibnd)) / a.dm[ibnd - 1];   // bdfrdx.f
        for (int i = 1; i <= margin; i++) {
            switch (kind) {
            case BD_FRE: Q(ibnd - i) = Q(ibnd); break;            // bdfrex.f
            case BD_SPP: Q(ibnd - i) = Q(ibnd - 1 + i); break;    // bdsppx.f
            case BD_SPN: Q(ibnd - i) = -Q(ibnd - 1 + i); break;   // bdspnx.f
            case BD_SMP: Q(ibnd - i) = Q(ibnd + i); break;        // bdsmpx.f
            case BD_SMN: Q(ibnd - i) = -Q(ibnd + i); break;       // bdsmnx.f
            case BD_CNS: Q(ibnd - i) = a.q0[f]; break;            // bdcnsx.f
            case BD_FRD: Q(ibnd - i) = Q(ibnd - i + 1) - dqq * a.dm[ibnd - i - 1]; break;
            default: break;
            }
        }
        if (kind == BD_SMN) Q(ibnd) = 0.0;
    } else {
        const int ibnd = n - margin;
        double dqq = 0.0;
        if (kind == BD_FRD) dqq = (Q(ibnd) - Q(ibnd - 1)) / a.dm[ibnd - 2];
        for (int i = 1; i <= margin; i++) {
            switch (kind) {
            case BD_FRE: Q(ibnd + i) = Q(ibnd); break;
            case BD_SPP: Q(ibnd + i) = Q(ibnd + 1 - i); break;
            case BD_SPN: Q(ibnd + i) = -Q(ibnd + 1 - i); break;
            case BD_SMP: Q(ibnd + i) = Q(ibnd - i); break;
            case BD_SMN: Q(ibnd + i) = -Q(ibnd - i); break;
            case BD_CNS: Q(ibnd + i) = a.q0[f]; break;
            case BD_FRD: Q(ibnd + i) = Q(ibnd + i - 1) + dqq * a.dm[ibnd + i - 2]; break;
            default: break;
            }
        }
        if (kind == BD_SMN) Q(ibnd) = 0.0;
    }
#undef Q
}

// Every field periodic on every face of a 1D/2D/3D box, all directions in
// one pass.  bdperx, then bdpery, then bdperz (each over every transverse
// index, the ghosts included) leave each ghost cell equal to the interior
// cell at its images along every axis, i -> i +- (n - 2m) for a ghost index
// (n - 2m >= m: one shift reaches the interior): each ghost cell is copied
// from that cell directly, the same values.  Sources are interior cells,
// which the pass never writes.  Wave units: (row with a ghost j or k, 64
// columns) copy the whole row with lanes along x; then 64 (row, ghost
// column) items of the rows whose j and k are interior.
struct PerFill {
    double *q[kMaxBndFields];
    int nf, ix, jx, kx, mx, my, mz;
    long long na, nb;   // units of the two kinds
};
CANS_DEV int per_img(int i, int n, int m) { return i <= m ? i + (n - 2 * m) : (i > n - m ? i - (n - 2 * m) : i); }

__global__ __launch_bounds__(256) void per_fill_kernel(PerFill a) {
    const int ix = a.ix, jx = a.jx, kx = a.kx, mx = a.mx, my = a.my, mz = a.mz;
    const long long nchunk = (ix + 63) / 64;
    const int lane = threadIdx.x & 63;
    const long long nw = (long long)gridDim.x * (blockDim.x >> 6);
    for (long long u = uni((long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)); u < a.na + a.nb; u += nw) {
        int i, j, k;
        bool on;
        if (u < a.na) {
            const long long g = u / nchunk;
            const int c = (int)(u - g * nchunk);
            const long long gz = 2LL * mz * jx;   // rows of the ghost z planes
            if (g < gz) {
                const int pk = (int)(g / jx);
                j = (int)(g - (long long)pk * jx) + 1;
                k = pk < mz ? pk + 1 : kx - 2 * mz + pk + 1;
            } else {                               // interior planes: their ghost y rows
                const long long h = g - gz;
                const int pj = (int)(h % (2 * my));
                k = mz + 1 + (int)(h / (2 * my));
                j = pj < my ? pj + 1 : jx - 2 * my + pj + 1;
            }
            i = c * 64 + lane + 1;
            on = i <= ix;
        } else {
            const long long v = (u - a.na) * 64 + lane;
            const long long nrow = (long long)(jx - 2 * my) * (kx - 2 * mz);
            on = v < nrow * 2 * mx;
            const int col = (int)(v % (2 * mx));
            const long long r = v / (2 * mx);
            j = my + 1 + (int)(r % (jx - 2 * my));
            k = mz + 1 + (int)(r / (jx - 2 * my));
            i = col < mx ? col + 1 : ix - 2 * mx + col + 1;
        }
        if (on) {
            const size_t dst = (size_t)(i - 1) + (size_t)ix * ((size_t)(j - 1) + (size_t)jx * (size_t)(k - 1));
            const size_t src = (size_t)(per_img(i, ix, mx) - 1) +
                               (size_t)ix * ((size_t)(per_img(j, jx, my) - 1) + (size_t)jx * (size_t)(per_img(k, kx, mz) - 1));
            for (int f = 0; f < a.nf; f++) a.q[f][dst] = a.q[f][src];
        }
    }
}

// ---------------------------------------------------------------- cfl
// 1D: cans1d/common/cfl_m.f:45-60, interior i=2..ix-1,
//     |vx| + sqrt(cs2+ca2), pi4i = 0.25/pi, onero = 1.0/ro.
// 2D: cans2d/common/cfl_m.f:44-62, all cells, min(dx,dy)/sqrt(v2+cs2+ca2),
//     pi4i = 1./pi/4.
// cfl_m3 (comp3): the 1D/2D formulas with bz in b2 and (2D) vz in v2.
// 3D: cans3d/common/cfl_m.f:39-58, all cells, min(dx,dy,dz)/sqrt(...),
//     onep4 = 0.25/pai with pai the REAL*4 acos(-1.0); spherical cfl_m_s
//     (trig set): min(dx, x dy, x sin(y) dz); 2D cfl_m_s (xr set): min(dx, x dy).
CANS_DEV double cfl_cell(const CflArgs &a, int i, int j, int k, size_t p) {
    const double gm = a.gm;
    if (a.dim == 1) {
        const double pi4i = 0.25 / kPi;
        const double onero = 1.0 / a.ro[p];
        double b2 = a.bx[p] * a.bx[p] + a.by[p] * a.by[p];
        if (a.comp3) b2 = b2 + a.bz[p] * a.bz[p];           // cans1d/common/cfl_m3.f:49
        const double ca2 = b2 * pi4i * onero;
        const double cs2 = gm * a.pr[p] * onero;
        const double dtcfl = a.dx[i - 1] / (__builtin_fabs(a.vx[p]) + __builtin_sqrt(cs2 + ca2));
        return a.safety * dtcfl;
    } else if (a.dim == 2) {
        const double pi4i = 1.0 / kPi / 4.0;
        double b2 = a.bx[p] * a.bx[p] + a.by[p] * a.by[p];
        if (a.comp3) b2 = b2 + a.bz[p] * a.bz[p];           // cans2d/common/cfl_m3.f:50-54
        const double ca2 = b2 * pi4i / a.ro[p];
        const double cs2 = gm * a.pr[p] / a.ro[p];
        double v2 = a.vx[p] * a.vx[p] + a.vy[p] * a.vy[p];
        if (a.comp3) v2 = v2 + a.vz[p] * a.vz[p];
        if (a.xr) {   // cans2d/common/cfl_m_s.f:54: min(dx, x dy)
            const double dtcfl = fminF(a.dx[i - 1], a.xr[i - 1] * a.dy[j - 1]) / __builtin_sqrt(v2 + cs2 + ca2);
            return a.safety * dtcfl;
        }
        const double dtcfl = fminF(a.dx[i - 1], a.dy[j - 1]) / __builtin_sqrt(v2 + cs2 + ca2);
        return a.safety * dtcfl;
    } else {
        const double pai = kPiF;            // pai = acos(-1.0), REAL*4
        const double pi4i = 0.25 / pai;     // onep4 = 0.25/pai
        const double b2 = a.bx[p] * a.bx[p] + a.by[p] * a.by[p] + a.bz[p] * a.bz[p];
        const double ca2 = b2 * pi4i / a.ro[p];
        const double cs2 = gm * a.pr[p] / a.ro[p];
        const double v2 = a.vx[p] * a.vx[p] + a.vy[p] * a.vy[p] + a.vz[p] * a.vz[p];
        if (a.trig) {   // cans3d/common/cfl_m_s.f:26-27: min(dx, x dy, x sin(y) dz)
            const double r = a.xr[i - 1];
            const double dtcfl = fminF(fminF(a.dx[i - 1], r * a.dy[j - 1]), r * a.trig[j - 1] * a.dz[k - 1])
                                 / __builtin_sqrt(v2 + cs2 + ca2);
            return a.safety * dtcfl;
        }
        const double dtcfl = fminF(fminF(a.dx[i - 1], a.dy[j - 1]), a.dz[k - 1]) / __builtin_sqrt(v2 + cs2 + ca2);
        return a.safety * dtcfl;
    }
}

CANS_DEV bool cfl_cell_in(const CflArgs &a, int i) { return a.dim != 1 || (i >= 2 && i <= a.ix - 1); }

__global__ __launch_bounds__(1024) void cfl_kernel(CflArgs a) {
    __shared__ double red[16];
    unsigned long long *key = a.key;
    if (a.clk) {
        if (!clock_live(a.clk, a.step)) return;   // uniform across the launch
        key = &a.clk->dtkey[a.step % 3];
    }
    double m = kR4_1e20;
    // waves take (row, 64-column chunk) items: the row indices are
    // wave-uniform, so no per-cell 64-bit division
    const int lane = threadIdx.x & 63;
    const int nch = (a.ix + 63) / 64;
    const long items = (long)nch * a.jx * a.kx;
    const long wstride = (long)gridDim.x * (blockDim.x >> 6);
    for (long it = uni((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6))); it < items; it += wstride) {
        const int ch = (int)(it % nch);
        const long row = it / nch;
        const int j = (int)(row % a.jx) + 1;
        const int k = (int)(row / a.jx) + 1;
        const int i = ch * 64 + lane + 1;
        if (i > a.ix || !cfl_cell_in(a, i)) continue;
        const size_t p = (size_t)(i - 1) + (size_t)row * (size_t)a.ix;
        const double q = cfl_cell(a, i, j, k, p);
        if (q < m) m = q;   // the reference's `if (dtq < dt)`: NaN never wins
    }
    block_min_to_global(m, red, key);
}

// first cell in the reference's loop order (i fastest) with dtq == dt
__global__ void cfl_locate_kernel(CflArgs a, double dt, unsigned long long *first) {
    const int lane = threadIdx.x & 63;
    const int nch = (a.ix + 63) / 64;
    const long items = (long)nch * a.jx * a.kx;
    const long wstride = (long)gridDim.x * (blockDim.x >> 6);
    for (long it = uni((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6))); it < items; it += wstride) {
        const int ch = (int)(it % nch);
        const long row = it / nch;
        const int j = (int)(row % a.jx) + 1;
        const int k = (int)(row / a.jx) + 1;
        const int i = ch * 64 + lane + 1;
        if (i > a.ix || !cfl_cell_in(a, i)) continue;
        const size_t p = (size_t)(i - 1) + (size_t)row * (size_t)a.ix;
        if (cfl_cell(a, i, j, k, p) == dt) atomicMin(first, (unsigned long long)p);
    }
}

__global__ void locate_arm(unsigned long long *first) { *first = ~0ull; }

__global__ void locate_finish(unsigned long long *first, int ix, int jx, int *loc) {
    const unsigned long long p = *first;
    if (p == ~0ull) { loc[0] = loc[1] = loc[2] = 0; return; }
    loc[0] = (int)(p % ix) + 1;
    loc[1] = (int)((p / ix) % jx) + 1;
    loc[2] = (int)(p / ((unsigned long long)ix * jx)) + 1;
}

// ---------------------------------------------------------------- copies
struct CopyArgs {
    double *dst[kMaxBndFields];
    const double *src[kMaxBndFields];
    int nf, ix, jx, kx, lo;
};
// copy the cells lo..n-lo+1 of every axis (the engines' update ranges)
__global__ void copy_interior_kernel(CopyArgs a) {
    const int ni = a.ix - 2 * (a.lo - 1), nj = a.jx > 1 ? a.jx - 2 * (a.lo - 1) : 1;
    const int nk = a.kx > 1 ? a.kx - 2 * (a.lo - 1) : 1;
    const size_t n = (size_t)ni * nj * nk;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n * a.nf; t += (size_t)gridDim.x * blockDim.x) {
        const int f = (int)(t / n);
        const size_t r = t - (size_t)f * n;
        const int i = (int)(r % ni) + a.lo;
        const int j = (int)((r / ni) % nj) + (a.jx > 1 ? a.lo : 1);
        const int k = (int)(r / ((size_t)ni * nj)) + (a.kx > 1 ? a.lo : 1);
        const size_t p = (size_t)(i - 1) + (size_t)a.ix * ((size_t)(j - 1) + (size_t)a.jx * (size_t)(k - 1));
        a.dst[f][p] = a.src[f][p];
    }
}

// ---------------------------------------------------------------- clock
__global__ void clock_reset_kernel(Clock *c, double t0, double tend, double dtmin, int nstop) {
    for (int s = 0; s < 3; s++) c->dtkey[s] = dkey(kR4_1e20);
    for (int s = 0; s < 2; s++) {
        c->t[s] = t0; c->dt_last[s] = 0.0; c->ns[s] = 0; c->nexec[s] = 0;
        c->stopped[s] = 0; c->merr[s] = 0;
    }
    c->tend = tend; c->dtmin = dtmin; c->nstop = nstop;
}
__global__ void arm_key_kernel(unsigned long long *k) { *k = dkey(kR4_1e20); }
__global__ void clock_step_kernel(Clock *c, int n) { (void)clock_enter(c, n, true); }
// the dt step n runs with (after its clock step), for a recorded trajectory
__global__ void clock_dt_log_kernel(const Clock *c, int n, double *out) { *out = c->dt_last[(n + 1) & 1]; }

}  // namespace

hipError_t launch_metric(const double *d, const double *dm, int n, Metric m, hipStream_t s) {
    hipLaunchKernelGGL(metric_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d, dm, n, m);
    return hipGetLastError();
}

hipError_t launch_bnd(const BndArgs &a, hipStream_t s) {
    const int n = a.dir == 0 ? a.ix : (a.dir == 1 ? a.jx : a.kx);
    const long nlines = (long)a.ix * a.jx * a.kx / n;
    const long total = nlines * a.nf;
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(bnd_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_bnd_periodic(double *const *q, int nf, int ix, int jx, int kx, int dim, int m, hipStream_t s) {
    if (nf < 1 || nf > kMaxBndFields || m < 1) return hipErrorInvalidValue;
    PerFill a{};
    for (int f = 0; f < nf; f++) a.q[f] = q[f];
    a.nf = nf; a.ix = ix; a.jx = jx; a.kx = kx;
    a.mx = m; a.my = dim >= 2 ? m : 0; a.mz = dim >= 3 ? m : 0;
    const int n[3] = {ix, jx, kx}, mm[3] = {a.mx, a.my, a.mz};
    for (int d = 0; d < 3; d++)
        if (mm[d] && n[d] - 2 * mm[d] < mm[d]) return hipErrorInvalidValue;   // one shift must reach the interior
    const long long nchunk = (ix + 63) / 64;
    a.na = (2LL * a.mz * jx + (long long)(kx - 2 * a.mz) * 2 * a.my) * nchunk;
    a.nb = ((long long)(jx - 2 * a.my) * (kx - 2 * a.mz) * 2 * a.mx + 63) / 64;
    long long blocks = (a.na + a.nb + 3) / 4;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) return hipSuccess;
    hipLaunchKernelGGL(per_fill_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_copy_interior(double *const *dst, const double *const *src, int nf,
                                int ix, int jx, int kx, int lo, hipStream_t s) {
    CopyArgs a{};
    for (int f = 0; f < nf; f++) { a.dst[f] = dst[f]; a.src[f] = src[f]; }
    a.nf = nf; a.ix = ix; a.jx = jx; a.kx = kx; a.lo = lo;
    const size_t n = (size_t)ix * jx * kx * nf;
    unsigned blocks = (unsigned)((n + 255) / 256);
    if (blocks > 65536u) blocks = 65536u;
    hipLaunchKernelGGL(copy_interior_kernel, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

// blocks over (row, 64-column chunk) items, one wave per item at a time
static unsigned cfl_blocks(const CflArgs &a, int waves_per_block, unsigned cap) {
    const size_t items = (size_t)((a.ix + 63) / 64) * (size_t)a.jx * (size_t)a.kx;
    const size_t b = (items + waves_per_block - 1) / waves_per_block;
    return (unsigned)(b < 1 ? 1 : (b > cap ? cap : b));
}

// 1024-thread blocks, at most 512: a quarter of the per-block dt-key
// atomics of 256-thread blocks at the same parallelism (2D 1024^2 pass
// 31 -> 18 us, 3D 256^3 243 -> 227 us, measured on one box)
hipError_t launch_cfl(const CflArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(cfl_kernel, dim3(cfl_blocks(a, 16, 512)), dim3(1024), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_cfl_locate(const CflArgs &a, double dt, int *loc_dev, hipStream_t s) {
    unsigned long long *first = reinterpret_cast<unsigned long long *>(loc_dev + 4);
    hipLaunchKernelGGL(locate_arm, dim3(1), dim3(1), 0, s, first);
    hipLaunchKernelGGL(cfl_locate_kernel, dim3(cfl_blocks(a, 4, 2048)), dim3(256), 0, s, a, dt, first);
    hipLaunchKernelGGL(locate_finish, dim3(1), dim3(1), 0, s, first, a.ix, a.jx, loc_dev);
    return hipGetLastError();
}

hipError_t launch_clock_reset(Clock *clk, double t0, double tend, double dtmin, int nstop, hipStream_t s) {
    hipLaunchKernelGGL(clock_reset_kernel, dim3(1), dim3(1), 0, s, clk, t0, tend, dtmin, nstop);
    return hipGetLastError();
}

hipError_t launch_arm_key(unsigned long long *key, hipStream_t s) {
    hipLaunchKernelGGL(arm_key_kernel, dim3(1), dim3(1), 0, s, key);
    return hipGetLastError();
}

hipError_t launch_clock_step(Clock *clk, int step, hipStream_t s) {
    hipLaunchKernelGGL(clock_step_kernel, dim3(1), dim3(1), 0, s, clk, step);
    return hipGetLastError();
}

hipError_t launch_clock_dt_log(const Clock *clk, int step, double *out, hipStream_t s) {
    hipLaunchKernelGGL(clock_dt_log_kernel, dim3(1), dim3(1), 0, s, clk, step, out);
    return hipGetLastError();
}

}  // namespace cans
