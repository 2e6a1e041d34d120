// cans_helpers.hip -- the helper routines of the CANS hot path as device
// calls with their own reference argument lists (mlwhalf, mlwfull, mlwartv,
// mlwsrch, mlwsrcf in 1D/2D/3D; roeflux_m and tvdminmod in 1D/2D/3D), for
// callers that compose their own engine from them.  The fused engines do not
// use these kernels; they exist so that every routine a CANS main program can
// link has a device counterpart.
//
// One thread per array element.  Each routine only reads its inputs and
// writes entries of its outputs no other thread writes (du, dan updated in
// place entry by entry), so a launch has no ordering hazards; entries the
// reference does not touch are left as they are.  Expressions keep the
// reference's association order (bit-exact FP64, no FMA contraction).
#include "cans_common.hpp"
#include "cans_roeflux.hpp"
#include "../../include/cans_gpu.h"

namespace cans {
int abi_fail(int code, const char *fmt, ...);
}

namespace {
using namespace cans;

#define HCHK(expr)                                                                                        \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) return abi_fail(CANS_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));      \
    } while (0)

inline dim3 grid_for(size_t n) { return dim3((unsigned)((n + 255) / 256)); }
__host__ __device__ inline size_t nn(int ix, int jx, int kx) { return (size_t)ix * jx * kx; }

struct Idx {   // 1-based (i,j,k) of a flat element p of an (ix,jx,kx) array
    int i, j, k;
};
CANS_DEV Idx idx3(size_t p, int ix, int jx) {
    return Idx{(int)(p % (size_t)ix) + 1, (int)((p / (size_t)ix) % (size_t)jx) + 1, (int)(p / ((size_t)ix * jx)) + 1};
}

// ------------------------------------------------------------------ mlwhalf
// cans1d/hdmlw/mlwhalf.f:33-46
__global__ void mlwhalf1_k(const double *u, double *un, double *du, double dt, const double *f, const double *dxi,
                           const double *dxim, int ix) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (i > ix) return;
    const int p = i - 1;
    if (i >= 2 && i <= ix - 1) du[p] = du[p] - 0.5 * dt * (0.5 * dxi[p] * (f[p + 1] - f[p - 1]));
    if (i <= ix - 1) {
        const double uh = 0.5 * (u[p + 1] + u[p]);
        const double dfdx = dxim[p] * (f[p + 1] - f[p]);
        un[p] = uh - dt * dfdx;
    }
}

// cans2d/hdmlw/mlwhalf.f:36-52
__global__ void mlwhalf2_k(const double *u, double *un, double *du, double dt, const double *fx, const double *dxi,
                           const double *dxim, int ix, const double *fy, const double *dyi, const double *dyim,
                           int jx) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nn(ix, jx, 1)) return;
    const Idx c = idx3(p, ix, jx);
    const size_t sy = ix;
    if (c.i >= 2 && c.i <= ix - 1 && c.j >= 2 && c.j <= jx - 1)
        du[p] = du[p] - 0.5 * dt * (0.5 * dxi[c.i - 1] * (fx[p + 1] - fx[p - 1]))
                - 0.5 * dt * (0.5 * dyi[c.j - 1] * (fy[p + sy] - fy[p - sy]));
    if (c.i <= ix - 1 && c.j <= jx - 1) {
        const double uh = 0.25 * (u[p + 1] + u[p] + u[p + 1 + sy] + u[p + sy]);
        const double dfdx = dxim[c.i - 1] * (fx[p + 1] - fx[p] + fx[p + 1 + sy] - fx[p + sy]) / 2.0;
        const double dfdy = dyim[c.j - 1] * (fy[p + sy] - fy[p] + fy[p + 1 + sy] - fy[p + 1]) / 2.0;
        un[p] = uh - dt * dfdx - dt * dfdy;
    }
}

// cans3d/hdmlw/mlwhalf.f:38-68
__global__ void mlwhalf3_k(const double *u, double *un, double *du, double dt, const double *fx, const double *dxi,
                           const double *dxim, int ix, const double *fy, const double *dyi, const double *dyim,
                           int jx, const double *fz, const double *dzi, const double *dzim, int kx) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nn(ix, jx, kx)) return;
    const Idx c = idx3(p, ix, jx);
    const size_t sy = ix, sz = (size_t)ix * jx;
    if (c.i >= 2 && c.i <= ix - 1 && c.j >= 2 && c.j <= jx - 1 && c.k >= 2 && c.k <= kx - 1)
        du[p] = du[p] - 0.5 * dt * (0.5 * dxi[c.i - 1] * (fx[p + 1] - fx[p - 1]))
                - 0.5 * dt * (0.5 * dyi[c.j - 1] * (fy[p + sy] - fy[p - sy]))
                - 0.5 * dt * (0.5 * dzi[c.k - 1] * (fz[p + sz] - fz[p - sz]));
    if (c.i <= ix - 1 && c.j <= jx - 1 && c.k <= kx - 1) {
        const size_t p100 = p + 1, p010 = p + sy, p110 = p + 1 + sy;
        const size_t p001 = p + sz, p101 = p001 + 1, p011 = p001 + sy, p111 = p011 + 1;
        const double uh = (u[p100] + u[p] + u[p110] + u[p010] + u[p101] + u[p001] + u[p111] + u[p011]) / 8.0;
        const double dfdx = (fx[p100] - fx[p] + fx[p110] - fx[p010] + fx[p101] - fx[p001] + fx[p111] - fx[p011])
                            / 4.0 * dxim[c.i - 1];
        const double dfdy = (fy[p010] - fy[p] + fy[p110] - fy[p100] + fy[p011] - fy[p001] + fy[p111] - fy[p101])
                            / 4.0 * dyim[c.j - 1];
        const double dfdz = (fz[p001] - fz[p] + fz[p101] - fz[p100] + fz[p011] - fz[p010] + fz[p111] - fz[p110])
                            / 4.0 * dzim[c.k - 1];
        un[p] = uh - dt * dfdx - dt * dfdy - dt * dfdz;
    }
}

// ------------------------------------------------------------------ mlwfull
// cans1d/hdmlw/mlwfull.f:30-33
__global__ void mlwfull1_k(double *du, double dt, const double *f, const double *dxi, int ix) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (i < 2 || i > ix - 1) return;
    const int p = i - 1;
    const double dfdx = dxi[p] * (f[p] - f[p - 1]);
    du[p] = du[p] - 0.5 * dt * dfdx;
}

// cans2d/hdmlw/mlwfull.f:31-41
__global__ void mlwfull2_k(double *du, double dt, const double *fx, const double *dxi, const double *ux0,
                           const double *ux1, int ix, const double *fy, const double *dyi, const double *uy0,
                           const double *uy1, int jx) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nn(ix, jx, 1)) return;
    const Idx c = idx3(p, ix, jx);
    if (c.i < 2 || c.i > ix - 1 || c.j < 2 || c.j > jx - 1) return;
    const size_t sy = ix;
    const int i = c.i - 1, j = c.j - 1;
    const double dfdx = dxi[i] * (uy0[j] * (fx[p - sy] - fx[p - 1 - sy]) + uy1[j] * (fx[p] - fx[p - 1]));
    const double dfdy = dyi[j] * (ux0[i] * (fy[p - 1] - fy[p - 1 - sy]) + ux1[i] * (fy[p] - fy[p - sy]));
    du[p] = du[p] - 0.5 * dt * dfdx - 0.5 * dt * dfdy;
}

// cans3d/hdmlw/mlwfull.f:34-55
__global__ void mlwfull3_k(double *du, double dt, const double *fx, const double *dxi, const double *ux0,
                           const double *ux1, int ix, const double *fy, const double *dyi, const double *uy0,
                           const double *uy1, int jx, const double *fz, const double *dzi, const double *uz0,
                           const double *uz1, int kx) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nn(ix, jx, kx)) return;
    const Idx c = idx3(p, ix, jx);
    if (c.i < 2 || c.i > ix - 1 || c.j < 2 || c.j > jx - 1 || c.k < 2 || c.k > kx - 1) return;
    const size_t sx = 1, sy = ix, sz = (size_t)ix * jx;
    const int i = c.i - 1, j = c.j - 1, k = c.k - 1;
    // neighbours at (i-1|i, j-1|j, k-1|k)
    const size_t q000 = p - sx - sy - sz, q100 = q000 + sx, q010 = q000 + sy, q110 = q010 + sx;
    const size_t q001 = q000 + sz, q101 = q001 + sx, q011 = q001 + sy, q111 = q011 + sx;
    const double dfdx = dxi[i] * (uz0[k] * (uy0[j] * (fx[q100] - fx[q000]) + uy1[j] * (fx[q110] - fx[q010]))
                                  + uz1[k] * (uy0[j] * (fx[q101] - fx[q001]) + uy1[j] * (fx[q111] - fx[q011])));
    const double dfdy = dyi[j] * (ux0[i] * (uz0[k] * (fy[q010] - fy[q000]) + uz1[k] * (fy[q011] - fy[q001]))
                                  + ux1[i] * (uz0[k] * (fy[q110] - fy[q100]) + uz1[k] * (fy[q111] - fy[q101])));
    const double dfdz = dzi[k] * (uy0[j] * (ux0[i] * (fz[q001] - fz[q000]) + ux1[i] * (fz[q101] - fz[q100]))
                                  + uy1[j] * (ux0[i] * (fz[q011] - fz[q010]) + ux1[i] * (fz[q111] - fz[q110])));
    du[p] = du[p] - 0.5 * dt * dfdx - 0.5 * dt * dfdy - 0.5 * dt * dfdz;
}

// ------------------------------------------------------------------ mlwartv
// cans1d/hdmlw/mlwartv.f:31-34
__global__ void mlwartv1_k(const double *u, double *du, double dt, const double *qx, const double *dxi,
                           const double *dxim, int ix) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (i < 2 || i > ix - 1) return;
    const int p = i - 1;
    du[p] = du[p] + dt * (dxi[p] * (qx[p] * dxim[p] * (u[p + 1] - u[p]) - qx[p - 1] * dxim[p - 1] * (u[p] - u[p - 1])));
}

// cans2d/hdmlw/mlwartv.f:33-41
__global__ void mlwartv2_k(const double *u, double *du, double dt, const double *qx, const double *dxi,
                           const double *dxim, int ix, const double *qy, const double *dyi, const double *dyim,
                           int jx) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nn(ix, jx, 1)) return;
    const Idx c = idx3(p, ix, jx);
    if (c.i < 2 || c.i > ix - 1 || c.j < 2 || c.j > jx - 1) return;
    const size_t sy = ix;
    const int i = c.i - 1, j = c.j - 1;
    du[p] = du[p] + dt * (dxi[i] * (qx[p] * dxim[i] * (u[p + 1] - u[p]) - qx[p - 1] * dxim[i - 1] * (u[p] - u[p - 1])))
            + dt * (dyi[j] * (qy[p] * dyim[j] * (u[p + sy] - u[p]) - qy[p - sy] * dyim[j - 1] * (u[p] - u[p - sy])));
}

// cans3d/hdmlw/mlwartv.f:36-48
__global__ void mlwartv3_k(const double *u, double *du, double dt, const double *qx, const double *dxi,
                           const double *dxim, int ix, const double *qy, const double *dyi, const double *dyim,
                           int jx, const double *qz, const double *dzi, const double *dzim, int kx) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nn(ix, jx, kx)) return;
    const Idx c = idx3(p, ix, jx);
    if (c.i < 2 || c.i > ix - 1 || c.j < 2 || c.j > jx - 1 || c.k < 2 || c.k > kx - 1) return;
    const size_t sy = ix, sz = (size_t)ix * jx;
    const int i = c.i - 1, j = c.j - 1, k = c.k - 1;
    du[p] = du[p] + dt * (dxi[i] * (qx[p] * dxim[i] * (u[p + 1] - u[p]) - qx[p - 1] * dxim[i - 1] * (u[p] - u[p - 1])))
            + dt * (dyi[j] * (qy[p] * dyim[j] * (u[p + sy] - u[p]) - qy[p - sy] * dyim[j - 1] * (u[p] - u[p - sy])))
            + dt * (dzi[k] * (qz[p] * dzim[k] * (u[p + sz] - u[p]) - qz[p - sz] * dzim[k - 1] * (u[p] - u[p - sz])));
}

// ------------------------------------------------------------------ sources
// cans1d/hdmlw/mlwsrch.f:32-41 (2D :32-47, 3D :32-47): source at the cells
// into du, averaged to the half-step points into un
__global__ void mlwsrch_k(double *un, double *du, double dt, const double *s, int ix, int jx, int kx, int dim) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nn(ix, jx, kx)) return;
    const Idx c = idx3(p, ix, jx);
    const size_t sy = ix, sz = (size_t)ix * jx;
    const bool in_i = c.i >= 2 && c.i <= ix - 1, in_j = dim < 2 || (c.j >= 2 && c.j <= jx - 1),
               in_k = dim < 3 || (c.k >= 2 && c.k <= kx - 1);
    if (in_i && in_j && in_k) du[p] = du[p] + 0.5 * dt * s[p];
    const bool h_i = c.i <= ix - 1, h_j = dim < 2 || c.j <= jx - 1, h_k = dim < 3 || c.k <= kx - 1;
    if (!(h_i && h_j && h_k)) return;
    double sh;
    if (dim == 1)
        sh = 0.5 * (s[p + 1] + s[p]);
    else if (dim == 2)
        sh = (s[p + 1] + s[p] + s[p + 1 + sy] + s[p + sy]) / 4.0;
    else
        sh = (s[p + 1] + s[p] + s[p + 1 + sy] + s[p + sy] + s[p + 1 + sz] + s[p + sz] + s[p + 1 + sy + sz]
              + s[p + sy + sz]) / 8.0;
    un[p] = un[p] + dt * sh;
}

// cans1d/hdmlw/mlwsrcf.f:31-34 (2D :33-40, 3D :33-46): the half-step source
// interpolated back to the cells with the metric weights
__global__ void mlwsrcf_k(double *du, double dt, const double *s, const double *ux0, const double *ux1,
                          const double *uy0, const double *uy1, const double *uz0, const double *uz1, int ix, int jx,
                          int kx, int dim) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nn(ix, jx, kx)) return;
    const Idx c = idx3(p, ix, jx);
    if (c.i < 2 || c.i > ix - 1) return;
    if (dim >= 2 && (c.j < 2 || c.j > jx - 1)) return;
    if (dim >= 3 && (c.k < 2 || c.k > kx - 1)) return;
    const size_t sy = ix, sz = (size_t)ix * jx;
    const int i = c.i - 1, j = c.j - 1, k = c.k - 1;
    double sh;
    if (dim == 1) {
        sh = ux0[i] * s[p - 1] + ux1[i] * s[p];
    } else if (dim == 2) {
        sh = uy0[j] * (ux0[i] * s[p - 1 - sy] + ux1[i] * s[p - sy]) + uy1[j] * (ux0[i] * s[p - 1] + ux1[i] * s[p]);
    } else {
        const size_t q = p - sz;
        sh = uz0[k] * (uy0[j] * (ux0[i] * s[q - 1 - sy] + ux1[i] * s[q - sy]) + uy1[j] * (ux0[i] * s[q - 1] + ux1[i] * s[q]))
             + uz1[k] * (uy0[j] * (ux0[i] * s[p - 1 - sy] + ux1[i] * s[p - sy])
                         + uy1[j] * (ux0[i] * s[p - 1] + ux1[i] * s[p]));
    }
    du[p] = du[p] + 0.5 * dt * sh;
}

// ------------------------------------------------------------------ Roe
// cans1d/hdroe/roeflux_m.f:48 (faces i=1..ix-1) and cans2d/hdroe/roeflux_m.f:16-17
// (j=1..jx-1, i=1..ix-1): w arrays hold the left state, then the right one,
// n = ix*jx apart
__global__ void roeflux12_k(double *fro, double *fee, double *frx, double *fry, double *fby, double gm,
                            const double *row, const double *prw, const double *vxw, const double *vyw,
                            const double *bxw, const double *byw, int ix, int jx) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = nn(ix, jx, 1);
    if (p >= n) return;
    const Idx c = idx3(p, ix, jx);
    if (c.i > ix - 1 || (jx > 1 && c.j > jx - 1)) return;
    const size_t q = p + n;
    double o[5];
    roeflux12<double>(gm, row[p], prw[p], vxw[p], vyw[p], bxw[p], byw[p], row[q], prw[q], vxw[q], vyw[q], bxw[q],
                      byw[q], o);
    fro[p] = o[0]; fee[p] = o[1]; frx[p] = o[2]; fry[p] = o[3]; fby[p] = o[4];
}

// cans3d/hdroe/roeflux_m.f:26-28 (k, j, i up to n-1), with the 2006 alphs term
__global__ void roeflux3_k(double *fro, double *fee, double *frx, double *fry, double *frz, double *fby, double *fbz,
                           double gm, const double *row, const double *prw, const double *vxw, const double *vyw,
                           const double *vzw, const double *bxw, const double *byw, const double *bzw, int ix, int jx,
                           int kx) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = nn(ix, jx, kx);
    if (p >= n) return;
    const Idx c = idx3(p, ix, jx);
    if (c.i > ix - 1 || c.j > jx - 1 || c.k > kx - 1) return;
    const size_t q = p + n;
    double o[7];
    roeflux3<double, true>(gm, row[p], prw[p], vxw[p], vyw[p], vzw[p], bxw[p], byw[p], bzw[p], row[q], prw[q],
                           vxw[q], vyw[q], vzw[q], bxw[q], byw[q], bzw[q], o);
    fro[p] = o[0]; fee[p] = o[1]; frx[p] = o[2]; fry[p] = o[3]; frz[p] = o[4]; fby[p] = o[5]; fbz[p] = o[6];
}

// flmt(a,b) = max(0, min(b*sign(1,a), |a|))*sign(1,a) of tvdminmod.f
CANS_DEV double flmt(double a, double b) {
    return fmaxF(0.0, fminF(b * signF(1.0, a), __builtin_fabs(a))) * signF(1.0, a);
}

// cans1d/hdroe/tvdminmod.f:33-40, cans2d/hdroe/tvdminmod.f:11-34 (every index
// 2..n-2, both directions) and cans3d/hdroe/tvdminmod.f:15-55 (flmt's
// arguments in the 3D order: minmod is symmetric, so the bits agree)
__global__ void tvdminmod_k(int mdir, const double *da, double *daw, int ix, int jx, int kx, int dim) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = nn(ix, jx, kx);
    if (p >= n) return;
    const Idx c = idx3(p, ix, jx);
    if (c.i < 2 || c.i > ix - 2) return;
    if (dim >= 2 && (c.j < 2 || c.j > jx - 2)) return;
    if (dim >= 3 && (c.k < 2 || c.k > kx - 2)) return;
    const size_t s = mdir == 1 ? 1 : (mdir == 2 ? (size_t)ix : (size_t)ix * jx);
    double w1, w2;
    if (dim < 3) {
        w1 = da[p] + 0.5 * flmt(da[p + s] - da[p], da[p] - da[p - s]);
        w2 = da[p + s] - 0.5 * flmt(da[p + s] - da[p], da[p + 2 * s] - da[p + s]);
    } else {
        w1 = da[p] + 0.5 * flmt(da[p] - da[p - s], da[p + s] - da[p]);
        w2 = da[p + s] - 0.5 * flmt(da[p + 2 * s] - da[p + s], da[p + s] - da[p]);
    }
    daw[p] = w1;
    daw[p + n] = w2;
}

int dims_ok(const char *who, int ix, int jx, int kx) {
    if (ix < 1 || jx < 1 || kx < 1) return abi_fail(CANS_ERR_ARG, "%s: ix=%d jx=%d kx=%d", who, ix, jx, kx);
    return 0;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- 1D
int cans1d_mlwhalf(const double *u, double *un, double *du, double dt, const double *f, const double *dxi,
                   const double *dxim, int ix, void *stream) {
    if (int e = dims_ok("cans1d_mlwhalf", ix, 1, 1)) return e;
    hipLaunchKernelGGL(mlwhalf1_k, grid_for(ix), dim3(256), 0, (hipStream_t)stream, u, un, du, dt, f, dxi, dxim, ix);
    HCHK(hipGetLastError());
    return 0;
}
int cans1d_mlwfull(double *du, double dt, const double *f, const double *dxi, int ix, void *stream) {
    if (int e = dims_ok("cans1d_mlwfull", ix, 1, 1)) return e;
    hipLaunchKernelGGL(mlwfull1_k, grid_for(ix), dim3(256), 0, (hipStream_t)stream, du, dt, f, dxi, ix);
    HCHK(hipGetLastError());
    return 0;
}
int cans1d_mlwartv(const double *u, double *du, double dt, const double *qx, const double *dxi, const double *dxim,
                   int ix, void *stream) {
    if (int e = dims_ok("cans1d_mlwartv", ix, 1, 1)) return e;
    hipLaunchKernelGGL(mlwartv1_k, grid_for(ix), dim3(256), 0, (hipStream_t)stream, u, du, dt, qx, dxi, dxim, ix);
    HCHK(hipGetLastError());
    return 0;
}
int cans1d_mlwsrch(double *un, double *du, double dt, const double *s, int ix, void *stream) {
    if (int e = dims_ok("cans1d_mlwsrch", ix, 1, 1)) return e;
    hipLaunchKernelGGL(mlwsrch_k, grid_for(ix), dim3(256), 0, (hipStream_t)stream, un, du, dt, s, ix, 1, 1, 1);
    HCHK(hipGetLastError());
    return 0;
}
int cans1d_mlwsrcf(double *du, double dt, const double *s, const double *ux0, const double *ux1, int ix,
                   void *stream) {
    if (int e = dims_ok("cans1d_mlwsrcf", ix, 1, 1)) return e;
    hipLaunchKernelGGL(mlwsrcf_k, grid_for(ix), dim3(256), 0, (hipStream_t)stream, du, dt, s, ux0, ux1,
                       (const double *)nullptr, (const double *)nullptr, (const double *)nullptr,
                       (const double *)nullptr, ix, 1, 1, 1);
    HCHK(hipGetLastError());
    return 0;
}
int cans1d_roeflux_m(double *fro, double *fee, double *frx, double *fry, double *fby, double gm, const double *row,
                     const double *prw, const double *vxw, const double *vyw, const double *bxw, const double *byw,
                     int ix, void *stream) {
    if (int e = dims_ok("cans1d_roeflux_m", ix, 1, 1)) return e;
    hipLaunchKernelGGL(roeflux12_k, grid_for(ix), dim3(256), 0, (hipStream_t)stream, fro, fee, frx, fry, fby, gm, row,
                       prw, vxw, vyw, bxw, byw, ix, 1);
    HCHK(hipGetLastError());
    return 0;
}
int cans1d_tvdminmod(const double *da, double *daw, int ix, void *stream) {
    if (int e = dims_ok("cans1d_tvdminmod", ix, 1, 1)) return e;
    hipLaunchKernelGGL(tvdminmod_k, grid_for(ix), dim3(256), 0, (hipStream_t)stream, 1, da, daw, ix, 1, 1, 1);
    HCHK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------- 2D
int cans2d_mlwhalf(const double *u, double *un, double *du, double dt, const double *fx, const double *dxi,
                   const double *dxim, int ix, const double *fy, const double *dyi, const double *dyim, int jx,
                   void *stream) {
    if (int e = dims_ok("cans2d_mlwhalf", ix, jx, 1)) return e;
    hipLaunchKernelGGL(mlwhalf2_k, grid_for(nn(ix, jx, 1)), dim3(256), 0, (hipStream_t)stream, u, un, du, dt, fx, dxi,
                       dxim, ix, fy, dyi, dyim, jx);
    HCHK(hipGetLastError());
    return 0;
}
int cans2d_mlwfull(double *du, double dt, const double *fx, const double *dxi, const double *ux0, const double *ux1,
                   int ix, const double *fy, const double *dyi, const double *uy0, const double *uy1, int jx,
                   void *stream) {
    if (int e = dims_ok("cans2d_mlwfull", ix, jx, 1)) return e;
    hipLaunchKernelGGL(mlwfull2_k, grid_for(nn(ix, jx, 1)), dim3(256), 0, (hipStream_t)stream, du, dt, fx, dxi, ux0,
                       ux1, ix, fy, dyi, uy0, uy1, jx);
    HCHK(hipGetLastError());
    return 0;
}
int cans2d_mlwartv(const double *u, double *du, double dt, const double *qx, const double *dxi, const double *dxim,
                   int ix, const double *qy, const double *dyi, const double *dyim, int jx, void *stream) {
    if (int e = dims_ok("cans2d_mlwartv", ix, jx, 1)) return e;
    hipLaunchKernelGGL(mlwartv2_k, grid_for(nn(ix, jx, 1)), dim3(256), 0, (hipStream_t)stream, u, du, dt, qx, dxi,
                       dxim, ix, qy, dyi, dyim, jx);
    HCHK(hipGetLastError());
    return 0;
}
int cans2d_mlwsrch(double *un, double *du, double dt, const double *s, int ix, int jx, void *stream) {
    if (int e = dims_ok("cans2d_mlwsrch", ix, jx, 1)) return e;
    hipLaunchKernelGGL(mlwsrch_k, grid_for(nn(ix, jx, 1)), dim3(256), 0, (hipStream_t)stream, un, du, dt, s, ix, jx,
                       1, 2);
    HCHK(hipGetLastError());
    return 0;
}
int cans2d_mlwsrcf(double *du, double dt, const double *s, const double *ux0, const double *ux1, int ix,
                   const double *uy0, const double *uy1, int jx, void *stream) {
    if (int e = dims_ok("cans2d_mlwsrcf", ix, jx, 1)) return e;
    hipLaunchKernelGGL(mlwsrcf_k, grid_for(nn(ix, jx, 1)), dim3(256), 0, (hipStream_t)stream, du, dt, s, ux0, ux1, uy0,
                       uy1, (const double *)nullptr, (const double *)nullptr, ix, jx, 1, 2);
    HCHK(hipGetLastError());
    return 0;
}
int cans2d_roeflux_m(double *fro, double *fee, double *frx, double *fry, double *fby, double gm, const double *row,
                     const double *prw, const double *vxw, const double *vyw, const double *bxw, const double *byw,
                     int ix, int jx, void *stream) {
    if (int e = dims_ok("cans2d_roeflux_m", ix, jx, 1)) return e;
    hipLaunchKernelGGL(roeflux12_k, grid_for(nn(ix, jx, 1)), dim3(256), 0, (hipStream_t)stream, fro, fee, frx, fry,
                       fby, gm, row, prw, vxw, vyw, bxw, byw, ix, jx);
    HCHK(hipGetLastError());
    return 0;
}
int cans2d_tvdminmod(int mdir, const double *da, double *daw, int ix, int jx, void *stream) {
    if (int e = dims_ok("cans2d_tvdminmod", ix, jx, 1)) return e;
    if (mdir != 1 && mdir != 2) return abi_fail(CANS_ERR_ARG, "cans2d_tvdminmod: mdir=%d", mdir);
    hipLaunchKernelGGL(tvdminmod_k, grid_for(nn(ix, jx, 1)), dim3(256), 0, (hipStream_t)stream, mdir, da, daw, ix, jx,
                       1, 2);
    HCHK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------- 3D
int cans3d_mlwhalf(const double *u, double *un, double *du, double dt, const double *fx, const double *dxi,
                   const double *dxim, int ix, const double *fy, const double *dyi, const double *dyim, int jx,
                   const double *fz, const double *dzi, const double *dzim, int kx, void *stream) {
    if (int e = dims_ok("cans3d_mlwhalf", ix, jx, kx)) return e;
    hipLaunchKernelGGL(mlwhalf3_k, grid_for(nn(ix, jx, kx)), dim3(256), 0, (hipStream_t)stream, u, un, du, dt, fx, dxi,
                       dxim, ix, fy, dyi, dyim, jx, fz, dzi, dzim, kx);
    HCHK(hipGetLastError());
    return 0;
}
int cans3d_mlwfull(double *du, double dt, const double *fx, const double *dxi, const double *ux0, const double *ux1,
                   int ix, const double *fy, const double *dyi, const double *uy0, const double *uy1, int jx,
                   const double *fz, const double *dzi, const double *uz0, const double *uz1, int kx, void *stream) {
    if (int e = dims_ok("cans3d_mlwfull", ix, jx, kx)) return e;
    hipLaunchKernelGGL(mlwfull3_k, grid_for(nn(ix, jx, kx)), dim3(256), 0, (hipStream_t)stream, du, dt, fx, dxi, ux0,
                       ux1, ix, fy, dyi, uy0, uy1, jx, fz, dzi, uz0, uz1, kx);
    HCHK(hipGetLastError());
    return 0;
}
int cans3d_mlwartv(const double *u, double *du, double dt, const double *qx, const double *dxi, const double *dxim,
                   int ix, const double *qy, const double *dyi, const double *dyim, int jx, const double *qz,
                   const double *dzi, const double *dzim, int kx, void *stream) {
    if (int e = dims_ok("cans3d_mlwartv", ix, jx, kx)) return e;
    hipLaunchKernelGGL(mlwartv3_k, grid_for(nn(ix, jx, kx)), dim3(256), 0, (hipStream_t)stream, u, du, dt, qx, dxi,
                       dxim, ix, qy, dyi, dyim, jx, qz, dzi, dzim, kx);
    HCHK(hipGetLastError());
    return 0;
}
int cans3d_mlwsrch(double *un, double *du, double dt, const double *s, int ix, int jx, int kx, void *stream) {
    if (int e = dims_ok("cans3d_mlwsrch", ix, jx, kx)) return e;
    hipLaunchKernelGGL(mlwsrch_k, grid_for(nn(ix, jx, kx)), dim3(256), 0, (hipStream_t)stream, un, du, dt, s, ix, jx,
                       kx, 3);
    HCHK(hipGetLastError());
    return 0;
}
int cans3d_mlwsrcf(double *du, double dt, const double *s, const double *ux0, const double *ux1, int ix,
                   const double *uy0, const double *uy1, int jx, const double *uz0, const double *uz1, int kx,
                   void *stream) {
    if (int e = dims_ok("cans3d_mlwsrcf", ix, jx, kx)) return e;
    hipLaunchKernelGGL(mlwsrcf_k, grid_for(nn(ix, jx, kx)), dim3(256), 0, (hipStream_t)stream, du, dt, s, ux0, ux1, uy0,
                       uy1, uz0, uz1, ix, jx, kx, 3);
    HCHK(hipGetLastError());
    return 0;
}
int cans3d_roeflux_m(double *fro, double *fee, double *frx, double *fry, double *frz, double *fby, double *fbz,
                     double gm, const double *row, const double *prw, const double *vxw, const double *vyw,
                     const double *vzw, const double *bxw, const double *byw, const double *bzw, int ix, int jx,
                     int kx, void *stream) {
    if (int e = dims_ok("cans3d_roeflux_m", ix, jx, kx)) return e;
    hipLaunchKernelGGL(roeflux3_k, grid_for(nn(ix, jx, kx)), dim3(256), 0, (hipStream_t)stream, fro, fee, frx, fry,
                       frz, fby, fbz, gm, row, prw, vxw, vyw, vzw, bxw, byw, bzw, ix, jx, kx);
    HCHK(hipGetLastError());
    return 0;
}
int cans3d_tvdminmod(int mdir, const double *da, double *daw, int ix, int jx, int kx, void *stream) {
    if (int e = dims_ok("cans3d_tvdminmod", ix, jx, kx)) return e;
    if (mdir < 1 || mdir > 3) return abi_fail(CANS_ERR_ARG, "cans3d_tvdminmod: mdir=%d", mdir);
    hipLaunchKernelGGL(tvdminmod_k, grid_for(nn(ix, jx, kx)), dim3(256), 0, (hipStream_t)stream, mdir, da, daw, ix,
                       jx, kx, 3);
    HCHK(hipGetLastError());
    return 0;
}

}  // extern "C"
