// cans_common.hpp -- shared device helpers for the CANS MHD kernels (gfx950).
//
// Arithmetic conventions that every kernel follows so that results are
// bit-identical to the reference Fortran compiled without FMA contraction:
//   * build with -ffp-contract=off; every expression keeps the reference's
//     left-to-right association, x**2 is x*x;
//   * MAX(a,b) = a > b ? a : b, MIN(a,b) = a < b ? a : b, SIGN(a,b) =
//     copysign(|a|, b) -- the semantics of the reference's compiled MAX, MIN
//     and SIGN, including NaN and signed-zero cases;
//   * REAL*4 literals that the reference mixes into double expressions are
//     promoted from float (1.0e-4 -> 9.99999974737875164e-05, 1.e20 ->
//     1.00000002004087734e+20, acos(-1.0) -> 3.14159274101257324);
//   * double division and sqrt are the correctly rounded IEEE operations.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define CANS_DEV __device__ __forceinline__

namespace cans {

constexpr double kPi = 3.141592653589793;           // acos(-1.0d0)
constexpr double kPiF = 3.14159274101257324;         // acos(-1.0), REAL*4
constexpr double kR4_1em4 = (double)1.0e-4f;         // REAL*4 1.0e-4
constexpr double kR4_1e20 = (double)1.e20f;          // REAL*4 1.e20

CANS_DEV double fmaxF(double a, double b) { return (a > b) ? a : b; }
CANS_DEV double fminF(double a, double b) { return (a < b) ? a : b; }
CANS_DEV double signF(double a, double b) { return __builtin_copysign(__builtin_fabs(a), b); }
CANS_DEV float fmaxF(float a, float b) { return (a > b) ? a : b; }
CANS_DEV float fminF(float a, float b) { return (a < b) ? a : b; }
CANS_DEV float signF(float a, float b) { return __builtin_copysignf(__builtin_fabsf(a), b); }
template <class T> CANS_DEV T sq(T a) { return a * a; }

// Block order across the XCDs.  The hardware deals the blocks of a launch
// to the 8 XCDs round-robin (block b to XCD b % 8), each with its own L2.
// A kernel that takes block b as the logical block (b % 8) * (grid / 8) +
// b / 8 of its linear order has each XCD work through one contiguous eighth
// of the grid, so the neighbouring rows its stencils share are fetched into
// one L2 instead of several.  Launch such kernels with xcd_grid() blocks (a
// multiple of 8; the extra blocks fall past the end and exit).
constexpr unsigned kXcds = 8;
CANS_DEV size_t xcd_first_thread() {
    const unsigned per = gridDim.x / kXcds;
    return (size_t)((blockIdx.x % kXcds) * per + blockIdx.x / kXcds) * blockDim.x;
}
inline unsigned xcd_grid(size_t nthreads, unsigned block = 256) {
    const size_t nb = (nthreads + block - 1) / block;
    return (unsigned)((nb + kXcds - 1) / kXcds * kXcds);
}

// Order-preserving map of a double onto uint64 so that an unsigned atomic
// min over the keys is a min over the doubles (NaNs are filtered before).
CANS_DEV uint64_t dkey(double d) {
    uint64_t u = (uint64_t)__double_as_longlong(d);
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
CANS_DEV double dunkey(uint64_t k) {
    uint64_t u = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)u);
}
inline uint64_t dkey_host(double d) {
    uint64_t u;
    __builtin_memcpy(&u, &d, 8);
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
inline double dunkey_host(uint64_t k) {
    uint64_t u = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
    double d;
    __builtin_memcpy(&d, &u, 8);
    return d;
}

// Wavefront (64 lanes) min of a double; NaN never wins because the caller
// only feeds values that passed `v < current`.
CANS_DEV double wave_min(double v) {
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(v, off, 64);
        v = (o < v) ? o : v;
    }
    return v;
}

// Block min through LDS, then one atomic per block on the order-preserving
// key, skipped when the key already holds a smaller or equal value (device-
// scope atomics on one address serialize: 2048 blocks that end together
// took ~20 us).  `smem` needs blockDim.x/64 doubles.
CANS_DEV void block_min_to_global(double v, double *smem, unsigned long long *gkey) {
    v = wave_min(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nw = (blockDim.x * blockDim.y * blockDim.z + 63) >> 6;
    if (lane == 0) smem[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0 && threadIdx.z == 0) {
        double m = smem[0];
        for (int w = 1; w < nw; w++) m = (smem[w] < m) ? smem[w] : m;
        const unsigned long long k = (unsigned long long)dkey(m);
        if (m < kR4_1e20 && k < __hip_atomic_load(gkey, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMin(gkey, k);
    }
}

// Blocks are dealt round-robin to the 8 XCDs of an MI355X.  Remap the
// hardware block id so that each XCD works on a contiguous run of tiles and
// tiles that share halo cells share one L2.
CANS_DEV int xcd_remap(int bid, int nblocks) {
    const int nx = 8;
    if (nblocks < nx * 2) return bid;
    const int per = nblocks / nx;          // full tiles per XCD
    const int full = per * nx;
    if (bid >= full) return bid;           // ragged tail keeps its id
    return (bid % nx) * per + bid / nx;
}

}  // namespace cans
