// cans_abi_spicule.hip -- extern "C" entry points of the UTRIP spicule model
// (declared in include/cans_gpu.h): mlw_m_bg, trqalfven, chkdav as single
// calls, and the whole time loop for many realizations.
#include <hip/hip_runtime.h>

#include "cans_common.hpp"
#include "cans_kernels.hpp"
#include "../../include/cans_gpu.h"

namespace cans {
int abi_fail(int code, const char *fmt, ...);
int abi_scratch(size_t bytes, void **out, hipStream_t s);
}  // namespace cans

using namespace cans;

#define HIPCHK(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return abi_fail(CANS_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                            __FILE__, __LINE__);                                             \
    } while (0)

namespace {

inline size_t al(size_t n) { return (n + 63) & ~size_t(63); }   // doubles, 512-byte aligned

// Park-Miller seed advance on the host: seed * 16807^n mod (2^31-1)
long long pm_skip(long long seed, long long n) {
    const unsigned long long M = 2147483647ull;
    unsigned long long r = (unsigned long long)seed % M, b = 16807ull;
    while (n > 0) {
        if (n & 1) r = r * b % M;
        b = b * b % M;
        n >>= 1;
    }
    return (long long)r;
}

// spicule workspace, in doubles: scratch state (5 fields of nrun*ix),
// metric (4 arrays per geometry set), draw multipliers (ix), counter
size_t sp_ws_doubles(int ix, int nrun, int nset) {
    return 5 * al((size_t)nrun * ix) + 4 * al((size_t)nset * ix) + al((size_t)ix) + 64;
}

}  // namespace

extern "C" {

int cans1d_mlw_m_bg(double *ro, double *pr, double *vx, double *vy, double *by, const double *bx,
                    const double *bxm, double dt, double qav, double gm, const double *gx, const double *gxm,
                    const double *sc, const double *dsc, const double *scm, const double *dscm, const double *rr,
                    const double *rrm, const double *drr, const double *drrm, const double *dx, const double *dxm,
                    int ix, void *stream) {
    if (ix < 3) return abi_fail(CANS_ERR_ARG, "cans1d_mlw_m_bg: ix=%d < 3", ix);
    const double *geo[SG_N] = {bx, bxm, gx, gxm, sc, dsc, scm, dscm, rr, rrm, drr, drrm, dx, dxm};
    for (int k = 0; k < SG_N; k++)
        if (!geo[k]) return abi_fail(CANS_ERR_ARG, "cans1d_mlw_m_bg: NULL geometry array %d", k);
    if (!ro || !pr || !vx || !vy || !by) return abi_fail(CANS_ERR_ARG, "cans1d_mlw_m_bg: NULL state array");
    hipStream_t s = (hipStream_t)stream;
    const size_t n = al((size_t)ix);
    void *w = nullptr;
    int rc = abi_scratch((5 + 4) * n * sizeof(double), &w, s);
    if (rc) return rc;
    double *ws = (double *)w;
    double *met[4] = {ws + 5 * n, ws + 6 * n, ws + 7 * n, ws + 8 * n};
    HIPCHK(launch_spicule_prep(dx, dxm, met, ix, 1, 0, nullptr, nullptr, s));
    MlwBgArgs a{};
    double *st[5] = {ro, pr, vx, vy, by};
    for (int q = 0; q < 5; q++) {
        a.in[q] = st[q];
        a.out[q] = ws + q * n;
    }
    for (int k = 0; k < SG_N; k++) a.geo[k] = geo[k];
    for (int k = 0; k < 4; k++) a.met[k] = met[k];
    a.dt = dt; a.qav = qav; a.gm = gm; a.ix = ix;
    HIPCHK(launch_mlw1d_bg(a, s));
    const double *src[5] = {a.out[0], a.out[1], a.out[2], a.out[3], a.out[4]};
    HIPCHK(launch_copy_interior(st, src, 5, ix, 1, 1, 2, s));
    return 0;
}

int cans1d_trqalfven(double *vy, double dt, const double *trq, const double *rr, int ix, int *mseed,
                     void *stream) {
    if (ix < 3 || !vy || !trq || !rr || !mseed) return abi_fail(CANS_ERR_ARG, "cans1d_trqalfven: bad argument");
    if (*mseed <= 0) return abi_fail(CANS_ERR_ARG, "cans1d_trqalfven: seed %d not in 1..2^31-2", *mseed);
    hipStream_t s = (hipStream_t)stream;
    void *w = nullptr;
    int rc = abi_scratch(al((size_t)ix) * sizeof(double), &w, s);
    if (rc) return rc;
    long long *pw = (long long *)w;
    double *nomet[4] = {nullptr, nullptr, nullptr, nullptr};
    HIPCHK(launch_spicule_prep(nullptr, nullptr, nomet, ix, 0, 0, pw, nullptr, s));
    HIPCHK(launch_trq1d(vy, dt, trq, rr, pw, *mseed, ix, s));
    *mseed = (int)pm_skip(*mseed, ix);
    return 0;
}

int cans1d_chkdav(int *n_events, double *da, double *vx, double floor_, int ix, void *stream) {
    if (ix < 1 || !da || !vx) return abi_fail(CANS_ERR_ARG, "cans1d_chkdav: bad argument");
    hipStream_t s = (hipStream_t)stream;
    void *w = nullptr;
    int rc = abi_scratch(64, &w, s);
    if (rc) return rc;
    int *cnt = (int *)w;
    HIPCHK(launch_chkdav1d(da, vx, floor_, ix, cnt, s));
    int c = 0;
    HIPCHK(hipMemcpyAsync(&c, cnt, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (n_events) *n_events = c;
    return 0;
}

size_t cans1d_spicule_ws_bytes(int ix, int nrun, int per_run_geometry) {
    if (ix < 3 || nrun < 1) return 0;
    return sp_ws_doubles(ix, nrun, per_run_geometry ? nrun : 1) * sizeof(double);
}

int cans1d_spicule_run(int nrun, int ix, int margin, double *const *state5, const double *const *geo14,
                       int per_run_geometry, const double *trq, double gm, double qav, double safety, double dtmin,
                       double floor_, double tend, int nstop, double *t, int *ns, int *merr, int *seed,
                       double dtout, double *snap, double *snap_t, int nsnap_max, int *nsnap, int block_threads,
                       int blocks, void *ws, void *stream) {
    if (nrun < 1 || !state5 || !geo14 || !trq || !t || !ns || !merr || !seed || !ws)
        return abi_fail(CANS_ERR_ARG, "cans1d_spicule_run: NULL argument or nrun < 1");
    if (margin < 1 || ix - 2 * margin < 2 * margin || ix - 2 * margin < 2)
        return abi_fail(CANS_ERR_ARG, "cans1d_spicule_run: ix=%d too small for margin=%d", ix, margin);
    for (int q = 0; q < 5; q++)
        if (!state5[q]) return abi_fail(CANS_ERR_ARG, "cans1d_spicule_run: NULL field %d", q);
    for (int k = 0; k < SG_N; k++)
        if (!geo14[k]) return abi_fail(CANS_ERR_ARG, "cans1d_spicule_run: NULL geometry array %d", k);
    if (snap && (!snap_t || nsnap_max < 1 || !(dtout > 0.0)))
        return abi_fail(CANS_ERR_ARG, "cans1d_spicule_run: snapshots need snap_t, nsnap_max >= 1, dtout > 0");
    if (block_threads != 64 && block_threads != 128 && block_threads != 256 && block_threads != 512 &&
        block_threads != 1024)
        return abi_fail(CANS_ERR_ARG, "cans1d_spicule_run: block_threads %d not in {64,...,1024}", block_threads);
    if (blocks < 1) blocks = nrun;
    hipStream_t s = (hipStream_t)stream;
    const int nset = per_run_geometry ? nrun : 1;
    const long long gstride = per_run_geometry ? ix : 0;
    const size_t N = al((size_t)nrun * ix), NG = al((size_t)nset * ix);
    double *w = (double *)ws;
    SpiculeArgs a{};
    a.nrun = nrun; a.ix = ix; a.margin = margin; a.gstride = gstride;
    for (int q = 0; q < 5; q++) {
        a.f[q] = state5[q];
        a.g[q] = w + q * N;
    }
    double *met[4];
    for (int k = 0; k < 4; k++) {
        met[k] = w + 5 * N + k * NG;
        a.met[k] = met[k];
    }
    long long *pw = (long long *)(w + 5 * N + 4 * NG);
    unsigned *next = (unsigned *)(w + 5 * N + 4 * NG + al((size_t)ix));
    for (int k = 0; k < SG_N; k++) a.geo[k] = geo14[k];
    a.trq = trq; a.pw = pw;
    a.gm = gm; a.qav = qav; a.safety = safety; a.dtmin = dtmin; a.floor = floor_; a.tend = tend;
    a.dtout = dtout; a.nstop = nstop;
    a.t = t; a.ns = ns; a.merr = merr; a.seed = seed;
    a.snap = snap; a.snap_t = snap_t; a.nsnap_max = nsnap_max; a.nsnap = nsnap;
    a.next = next;
    HIPCHK(launch_spicule_prep(geo14[SG_DX], geo14[SG_DXM], met, ix, nset, gstride, pw, next, s));
    HIPCHK(launch_spicule(a, block_threads, blocks, s));
    return 0;
}

}  // extern "C"
