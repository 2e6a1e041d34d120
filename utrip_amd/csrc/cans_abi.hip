// cans_abi.hip -- the extern "C" boundary of libcans_gpu.so.
//
// Three layers, all plain pointers and sizes (declared in include/cans_gpu.h):
//  * device-array entry points with the CANS argument lists (one call = one
//    reference call, in place on the caller's arrays);
//  * the resident time loop (device clock + multi-step runs, no host round
//    trip per step);
//  * host-array mapping used by the Fortran drop-in wrappers.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <cmath>
#include <atomic>
#include <map>
#include <mutex>
#include <vector>

#include "cans_common.hpp"
#include "cans_kernels.hpp"
#include "cans_roe.hpp"
#include "../../include/cans_gpu.h"

using namespace cans;

namespace cans {
thread_local char g_err[512] = "";
int abi_fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}
int abi_scratch(size_t bytes, void **out, hipStream_t s);
int host_trig(const double *y, const double *ym, int jx, double *dev, hipStream_t s);
}  // namespace cans

namespace {

std::mutex g_mu;
#define fail cans::abi_fail
#define HIPCHK(expr)                                                                    \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(CANS_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                            \
    } while (0)

// Scratch device memory of the drop-in entry points: one pool per (device,
// stream), grown to the largest request and kept until
// cans_gpu_release_all.  Calls queued on one stream run in order and may
// share their pool; calls on different streams get different pools, so they
// never overwrite each other's workspace.  Host threads that enqueue on the
// same stream at once still race on it (as they would on the stream).
struct Pool {
    void *ptr = nullptr;
    size_t bytes = 0;
};
std::map<std::pair<int, hipStream_t>, Pool> g_pool;

int pool_scratch(size_t bytes, void **out, hipStream_t stream) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_mu);
    Pool &p = g_pool[{dev, stream}];
    if (p.bytes < bytes) {
        if (p.ptr) {
            HIPCHK(hipDeviceSynchronize());
            HIPCHK(hipFree(p.ptr));
            p.ptr = nullptr;
            p.bytes = 0;
        }
        HIPCHK(hipMalloc(&p.ptr, bytes));
        p.bytes = bytes;
    }
    *out = p.ptr;
    return 0;
}

inline size_t al(size_t n) { return (n + 63) & ~size_t(63); }  // doubles, 512-byte aligned

}  // namespace

int cans::abi_scratch(size_t bytes, void **out, hipStream_t s) { return pool_scratch(bytes, out, s); }

namespace {
std::atomic<int> g_march_waves{32768};
std::atomic<int> g_mlw3d_path{1};
}
int cans::march_waves() { return g_march_waves.load(); }
int cans::mlw3d_path() { return g_mlw3d_path.load(); }

// sin, cos, tan of the colatitudes y (and ym) on the host -- the C library's
// functions, as the reference's compiled intrinsics -- into dev[0..3jx)
// (y) and dev[3jx..6jx) (ym, when given): the tables of the spherical engine.
int cans::host_trig(const double *y, const double *ym, int jx, double *dev, hipStream_t s) {
    std::vector<double> h(2 * (size_t)jx), t(6 * (size_t)jx, 0.0);
    HIPCHK(hipMemcpyAsync(h.data(), y, jx * sizeof(double), hipMemcpyDeviceToHost, s));
    if (ym) HIPCHK(hipMemcpyAsync(h.data() + jx, ym, jx * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int side = 0; side < (ym ? 2 : 1); side++)
        for (int j = 0; j < jx; j++) {
            const double v = h[(size_t)side * jx + j];
            t[(size_t)(3 * side) * jx + j] = std::sin(v);
            t[(size_t)(3 * side + 1) * jx + j] = std::cos(v);
            t[(size_t)(3 * side + 2) * jx + j] = std::tan(v);
        }
    HIPCHK(hipMemcpyAsync(dev, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

extern "C" {

const char *cans_gpu_last_error(void) { return cans::g_err; }

int cans_gpu_init(int device) {
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(CANS_ERR_ARG, "device %d of %d", device, n);
    HIPCHK(hipSetDevice(device));
    return 0;
}

int cans_gpu_sync(void *stream) {
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    return 0;
}

int cans_gpu_set_march_waves(int waves) {
    if (waves < 1) return fail(CANS_ERR_ARG, "march waves %d < 1", waves);
    g_march_waves.store(waves);
    return 0;
}

int cans_gpu_set_mlw3d_path(int path) {
    if (path != 0 && path != 1) return fail(CANS_ERR_ARG, "mlw3d path %d (0 or 1)", path);
    g_mlw3d_path.store(path);
    return 0;
}

// ====================================================================== 1D
// One mlw_m 1D call on the workspace ws of 9*al(ix) doubles: the five
// updated fields, then 1/dx, 1/dxm, ux0, ux1.
static int mlw1d_on_ws(double *ro, double *pr, double *vx, double *vy, double *by, const double *bx,
                       const double *bxm, double dt, double qav, double gm, const double *dx, const double *dxm,
                       int ix, double *ws, hipStream_t s) {
    const size_t n = al(ix);
    Metric mx{ws + 5 * n, ws + 6 * n, ws + 7 * n, ws + 8 * n};
    HIPCHK(launch_metric(dx, dxm, ix, mx, s));
    Mlw1dArgs a{};
    a.ro = ro; a.pr = pr; a.vx = vx; a.vy = vy; a.by = by;
    a.ro_o = ws; a.pr_o = ws + n; a.vx_o = ws + 2 * n; a.vy_o = ws + 3 * n; a.by_o = ws + 4 * n;
    a.bx = bx; a.bxm = bxm; a.dx = dx; a.dxm = dxm; a.mx = mx;
    a.qav = qav; a.gm = gm; a.dt = dt; a.clk = nullptr; a.ix = ix;
    HIPCHK(launch_mlw1d(a, s));
    double *dst[5] = {ro, pr, vx, vy, by};
    const double *src[5] = {a.ro_o, a.pr_o, a.vx_o, a.vy_o, a.by_o};
    HIPCHK(launch_copy_interior(dst, src, 5, ix, 1, 1, 2, s));
    return 0;
}

int cans1d_mlw_m(double *ro, double *pr, double *vx, double *vy, double *by,
                 const double *bx, const double *bxm, double dt, double qav, double gm,
                 const double *dx, const double *dxm, int ix, void *stream) {
    if (ix < 3) return fail(CANS_ERR_ARG, "cans1d_mlw_m: ix=%d < 3", ix);
    void *w = nullptr;
    int rc = pool_scratch(9 * al(ix) * sizeof(double), &w, (hipStream_t)stream);
    if (rc) return rc;
    return mlw1d_on_ws(ro, pr, vx, vy, by, bx, bxm, dt, qav, gm, dx, dxm, ix, (double *)w, (hipStream_t)stream);
}

}  // extern "C"

// Workspace of cans1d_mlw_m_noalloc: a file-scope device array, one copy per
// GPU, held until the process ends.  Host code never passes `scratch` itself
// (on the host it names a stand-in, not device memory): it passes the address
// hipGetSymbolAddress returns for the current device, looked up once per
// device and kept in g_noalloc_ptr.
static __device__ __align__(256) double scratch[9 * CANS_NOALLOC_MAX_IX];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
double *g_noalloc_ptr[kMaxDevices] = {};

int noalloc_scratch(double **out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(CANS_ERR_ARG, "device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(CANS_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL on device %d", dev);
        g_noalloc_ptr[dev] = (double *)p;
    }
    *out = g_noalloc_ptr[dev];
    return 0;
}
}  // namespace

extern "C" {

int cans1d_mlw_m_noalloc(double *ro, double *pr, double *vx, double *vy, double *by, const double *bx,
                         const double *bxm, double dt, double qav, double gm, const double *dx, const double *dxm,
                         int ix, void *stream) {
    if (ix < 3) return fail(CANS_ERR_ARG, "cans1d_mlw_m_noalloc: ix=%d < 3", ix);
    if (ix > CANS_NOALLOC_MAX_IX)
        return fail(CANS_ERR_ARG, "cans1d_mlw_m_noalloc: ix=%d exceeds the static workspace (ix <= %d)", ix,
                    CANS_NOALLOC_MAX_IX);
    double *ws = nullptr;
    int rc = noalloc_scratch(&ws);
    if (rc) return rc;
    return mlw1d_on_ws(ro, pr, vx, vy, by, bx, bxm, dt, qav, gm, dx, dxm, ix, ws, (hipStream_t)stream);
}

static int cfl_common(CflArgs &a, double *dt, double dtmin, int *merr, hipStream_t s) {
    void *w = nullptr;
    int rc = pool_scratch(64, &w, s);
    if (rc) return rc;
    unsigned long long *key = (unsigned long long *)w;
    int *loc = (int *)((char *)w + 8);
    HIPCHK(launch_arm_key(key, s));
    a.key = key;
    a.clk = nullptr;
    HIPCHK(launch_cfl(a, s));
    unsigned long long k = 0;
    HIPCHK(hipMemcpyAsync(&k, key, 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    *dt = dunkey_host(k);
    *merr = 0;
    if (*dt < dtmin) {   // cfl_m.f:64-71
        *merr = 9001;
        int l[3] = {0, 0, 0};
        HIPCHK(launch_cfl_locate(a, *dt, loc, s));
        HIPCHK(hipMemcpyAsync(l, loc, sizeof(l), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        printf("  ### stop due to small dt, less than dtmin ###\n");
        if (a.dim == 1) printf("   dt = %10.3E  < %10.3E @ i =%5d\n", *dt, dtmin, l[0]);
        else if (a.dim == 2) printf("   dt = %10.3E  < %10.3E @ i =%5d j = %5d\n", *dt, dtmin, l[0], l[1]);
        else printf("   dt = %10.3E  < %10.3E @ i =%5d j = %5d k = %5d\n", *dt, dtmin, l[0], l[1], l[2]);
        fflush(stdout);
    }
    return 0;
}

int cans1d_cfl_m(double *dt, double safety, double dtmin, int *merr, double gm,
                 const double *bx, const double *ro, const double *pr, const double *vx,
                 const double *vy, const double *by, const double *dx, int ix, void *stream) {
    if (ix < 3) return fail(CANS_ERR_ARG, "cans1d_cfl_m: ix=%d < 3", ix);
    CflArgs a{};
    a.ro = ro; a.pr = pr; a.vx = vx; a.vy = vy; a.bx = bx; a.by = by; a.dx = dx;
    a.safety = safety; a.gm = gm; a.ix = ix; a.jx = 1; a.kx = 1; a.dim = 1;
    return cfl_common(a, dt, dtmin, merr, (hipStream_t)stream);
}

int cans1d_cfl_m3(double *dt, double safety, double dtmin, int *merr, double gm, const double *bx,
                  const double *ro, const double *pr, const double *vx, const double *vy, const double *by,
                  const double *vz, const double *bz, const double *dx, int ix, void *stream) {
    if (ix < 3) return fail(CANS_ERR_ARG, "cans1d_cfl_m3: ix=%d < 3", ix);
    CflArgs a{};
    a.ro = ro; a.pr = pr; a.vx = vx; a.vy = vy; a.vz = vz; a.bx = bx; a.by = by; a.bz = bz; a.dx = dx;
    a.safety = safety; a.gm = gm; a.ix = ix; a.jx = 1; a.kx = 1; a.dim = 1; a.comp3 = 1;
    return cfl_common(a, dt, dtmin, merr, (hipStream_t)stream);
}

static int bnd_common(int dir, int kind, int mbnd, int margin, double *qq, double q0,
                      const double *dm, int ix, int jx, int kx, void *stream) {
    const int n = dir == 0 ? ix : (dir == 1 ? jx : kx);
    if (kind < BD_FRE || kind > BD_PER) return fail(CANS_ERR_ARG, "bnd: kind %d", kind);
    if (kind == BD_PER) {
        if (mbnd < 0 || margin < 0 || mbnd + margin > n) return fail(CANS_ERR_ARG, "bdper: margins %d,%d n=%d", mbnd, margin, n);
    } else {
        if (mbnd != 0 && mbnd != 1) return fail(CANS_ERR_ARG, "bnd: mbnd=%d", mbnd);
        if (margin < 0 || 2 * margin + 1 > n) return fail(CANS_ERR_ARG, "bnd: margin %d n=%d", margin, n);
        if (kind == BD_FRD && !dm) return fail(CANS_ERR_ARG, "bdfrd: dm is NULL");
    }
    BndArgs a{};
    a.q[0] = qq; a.q0[0] = q0; a.kind[0] = kind; a.nf = 1;
    a.dir = dir; a.mbnd = mbnd; a.margin = margin; a.dm = dm;
    a.ix = ix; a.jx = jx; a.kx = kx;
    HIPCHK(launch_bnd(a, (hipStream_t)stream));
    return 0;
}

int cans1d_bnd(int kind, int mbnd, int margin, double *qq, double q0, const double *dxm,
               int ix, void *stream) {
    return bnd_common(0, kind, mbnd, margin, qq, q0, dxm, ix, 1, 1, stream);
}

// ====================================================================== 2D
int cans2d_mlw_m(double *ro, double *pr, double *vx, double *vy, double *bx, double *by,
                 double *az, double dt, double qav, double gm, const double *dx,
                 const double *dxm, int ix, const double *dy, const double *dym, int jx,
                 void *stream) {
    if (ix < 3 || jx < 3) return fail(CANS_ERR_ARG, "cans2d_mlw_m: ix=%d jx=%d", ix, jx);
    hipStream_t s = (hipStream_t)stream;
    const size_t n = al((size_t)ix * jx), mxn = al(ix), myn = al(jx);
    void *w = nullptr;
    int rc = pool_scratch((7 * n + 4 * mxn + 4 * myn) * sizeof(double), &w, s);
    if (rc) return rc;
    double *ws = (double *)w;
    double *o[7];
    for (int f = 0; f < 7; f++) o[f] = ws + f * n;
    double *m = ws + 7 * n;
    Metric mx{m, m + mxn, m + 2 * mxn, m + 3 * mxn};
    m += 4 * mxn;
    Metric my{m, m + myn, m + 2 * myn, m + 3 * myn};
    HIPCHK(launch_metric(dx, dxm, ix, mx, s));
    HIPCHK(launch_metric(dy, dym, jx, my, s));
    Mlw2dArgs a{};
    a.ro = ro; a.pr = pr; a.vx = vx; a.vy = vy; a.bx = bx; a.by = by; a.az = az;
    a.ro_o = o[0]; a.pr_o = o[1]; a.vx_o = o[2]; a.vy_o = o[3]; a.bx_o = o[4]; a.by_o = o[5]; a.az_o = o[6];
    a.dx = dx; a.dxm = dxm; a.dy = dy; a.dym = dym; a.mx = mx; a.my = my;
    a.qav = qav; a.gm = gm; a.dt = dt; a.clk = nullptr; a.ix = ix; a.jx = jx;
    a.per_margin = 0; a.safety = 0.0;
    HIPCHK(launch_mlw2d(a, s));
    double *dst[7] = {ro, pr, vx, vy, bx, by, az};
    const double *src[7] = {o[0], o[1], o[2], o[3], o[4], o[5], o[6]};
    HIPCHK(launch_copy_interior(dst, src, 7, ix, jx, 1, 2, s));
    return 0;
}

int cans2d_cfl_m(double *dt, double safety, double dtmin, int *merr, double gm,
                 const double *ro, const double *pr, const double *vx, const double *vy,
                 const double *bx, const double *by, const double *dx, int ix,
                 const double *dy, int jx, void *stream) {
    CflArgs a{};
    a.ro = ro; a.pr = pr; a.vx = vx; a.vy = vy; a.bx = bx; a.by = by; a.dx = dx; a.dy = dy;
    a.safety = safety; a.gm = gm; a.ix = ix; a.jx = jx; a.kx = 1; a.dim = 2;
    return cfl_common(a, dt, dtmin, merr, (hipStream_t)stream);
}

int cans2d_cfl_m_s(double *dt, double safety, double dtmin, int *merr, double gm, const double *ro,
                   const double *pr, const double *vx, const double *vy, const double *bx, const double *by,
                   const double *x, const double *dx, int ix, const double *dy, int jx, void *stream) {
    if (!x) return fail(CANS_ERR_ARG, "cans2d_cfl_m_s: NULL x");
    CflArgs a{};
    a.ro = ro; a.pr = pr; a.vx = vx; a.vy = vy; a.bx = bx; a.by = by; a.dx = dx; a.dy = dy; a.xr = x;
    a.safety = safety; a.gm = gm; a.ix = ix; a.jx = jx; a.kx = 1; a.dim = 2;
    return cfl_common(a, dt, dtmin, merr, (hipStream_t)stream);
}

int cans2d_cfl_m3(double *dt, double safety, double dtmin, int *merr, double gm, const double *ro,
                  const double *pr, const double *vx, const double *vy, const double *vz, const double *bx,
                  const double *by, const double *bz, const double *dx, int ix, const double *dy, int jx,
                  void *stream) {
    CflArgs a{};
    a.ro = ro; a.pr = pr; a.vx = vx; a.vy = vy; a.vz = vz; a.bx = bx; a.by = by; a.bz = bz; a.dx = dx; a.dy = dy;
    a.safety = safety; a.gm = gm; a.ix = ix; a.jx = jx; a.kx = 1; a.dim = 2; a.comp3 = 1;
    return cfl_common(a, dt, dtmin, merr, (hipStream_t)stream);
}

int cans2d_bnd(int dir, int kind, int mbnd, int margin, double *qq, double q0,
               const double *dm, int ix, int jx, void *stream) {
    if (dir < 0 || dir > 1) return fail(CANS_ERR_ARG, "cans2d_bnd: dir %d", dir);
    return bnd_common(dir, kind, mbnd, margin, qq, q0, dm, ix, jx, 1, stream);
}

// ====================================================================== 3D bnd/cfl
int cans3d_cfl_m(double *dt, double safety, double dtmin, int *merr, double gm,
                 const double *ro, const double *pr, const double *vx, const double *vy,
                 const double *vz, const double *bx, const double *by, const double *bz,
                 const double *dx, int ix, const double *dy, int jx, const double *dz, int kx,
                 void *stream) {
    CflArgs a{};
    a.ro = ro; a.pr = pr; a.vx = vx; a.vy = vy; a.vz = vz; a.bx = bx; a.by = by; a.bz = bz;
    a.dx = dx; a.dy = dy; a.dz = dz;
    a.safety = safety; a.gm = gm; a.ix = ix; a.jx = jx; a.kx = kx; a.dim = 3;
    return cfl_common(a, dt, dtmin, merr, (hipStream_t)stream);
}

int cans3d_cfl_m_s(double *dt, double safety, double dtmin, int *merr, double gm,
                   const double *ro, const double *pr, const double *vx, const double *vy,
                   const double *vz, const double *bx, const double *by, const double *bz,
                   const double *x, const double *dx, int ix, const double *y, const double *dy, int jx,
                   const double *dz, int kx, void *stream) {
    if (!x || !y) return fail(CANS_ERR_ARG, "cans3d_cfl_m_s: NULL x or y");
    hipStream_t s = (hipStream_t)stream;
    double *tab = nullptr;
    if (int rc = pool_scratch(64 + 6 * (size_t)jx * sizeof(double), (void **)&tab, s)) return rc;
    tab += 8;                                   // the first 64 bytes hold cfl_common's key
    if (int rc = cans::host_trig(y, nullptr, jx, tab, s)) return rc;
    CflArgs a{};
    a.ro = ro; a.pr = pr; a.vx = vx; a.vy = vy; a.vz = vz; a.bx = bx; a.by = by; a.bz = bz;
    a.dx = dx; a.dy = dy; a.dz = dz; a.xr = x; a.trig = tab;
    a.safety = safety; a.gm = gm; a.ix = ix; a.jx = jx; a.kx = kx; a.dim = 3;
    return cfl_common(a, dt, dtmin, merr, s);
}

int cans3d_bnd(int dir, int kind, int mbnd, int margin, double *qq, double q0,
               const double *dm, int ix, int jx, int kx, void *stream) {
    if (dir < 0 || dir > 2) return fail(CANS_ERR_ARG, "cans3d_bnd: dir %d", dir);
    return bnd_common(dir, kind, mbnd, margin, qq, q0, dm, ix, jx, kx, stream);
}

// ====================================================================== clock
int cans_clock_create(void **clk) {
    if (!clk) return fail(CANS_ERR_ARG, "clk is NULL");
    HIPCHK(hipMalloc(clk, sizeof(Clock)));
    return 0;
}
int cans_clock_destroy(void *clk) {
    if (clk) HIPCHK(hipFree(clk));
    return 0;
}
int cans_clock_reset(void *clk, double t0, double tend, int nstop, double dtmin, void *stream) {
    if (!clk) return fail(CANS_ERR_ARG, "clk is NULL");
    HIPCHK(launch_clock_reset((Clock *)clk, t0, tend, dtmin, nstop, (hipStream_t)stream));
    return 0;
}
int cans_clock_read(const void *clk, int nsteps_issued, double *t, int *ns, int *nexec,
                    int *merr, double *dt_last, void *stream) {
    if (!clk) return fail(CANS_ERR_ARG, "clk is NULL");
    Clock c;
    HIPCHK(hipMemcpyAsync(&c, clk, sizeof(Clock), hipMemcpyDeviceToHost, (hipStream_t)stream));
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    const int s = nsteps_issued & 1;
    if (t) *t = c.t[s];
    if (ns) *ns = c.ns[s];
    if (nexec) *nexec = c.nexec[s];
    if (merr) *merr = c.merr[s];
    if (dt_last) *dt_last = c.dt_last[s];
    return 0;
}

// 2D MLW resident run: state alternates a (even steps) -> b -> a ...
int cans2d_mlw_run(double *const *a7, double *const *b7, void *clk, int step0, int nsteps,
                   double qav, double gm, double safety, const double *dx, const double *dxm,
                   int ix, const double *dy, const double *dym, int jx, const int *bc,
                   int margin, double *metric_ws, void *const *events, void *stream) {
    if (!clk || !a7 || !b7 || !metric_ws) return fail(CANS_ERR_ARG, "cans2d_mlw_run: NULL argument");
    if (ix < 3 || jx < 3 || margin < 1) return fail(CANS_ERR_ARG, "cans2d_mlw_run: ix=%d jx=%d margin=%d", ix, jx, margin);
    hipStream_t s = (hipStream_t)stream;
    Clock *c = (Clock *)clk;
    const size_t mxn = al(ix), myn = al(jx);
    double *m = metric_ws;
    Metric mx{m, m + mxn, m + 2 * mxn, m + 3 * mxn};
    m += 4 * mxn;
    Metric my{m, m + myn, m + 2 * myn, m + 3 * myn};
    if (step0 == 0) {   // the first call of a run fills the metric workspace
        HIPCHK(launch_metric(dx, dxm, ix, mx, s));
        HIPCHK(launch_metric(dy, dym, jx, my, s));
    }
    const bool periodic = (bc == nullptr);
    if (periodic && (ix - 2 * margin < margin || jx - 2 * margin < margin))
        return fail(CANS_ERR_ARG, "cans2d_mlw_run: periodic box needs ix-2*margin >= margin (ix=%d jx=%d margin=%d)", ix, jx, margin);
    for (int n = step0; n < step0 + nsteps; n++) {
        double *const *in = (n & 1) ? b7 : a7;
        double *const *out = (n & 1) ? a7 : b7;
        if (!periodic || n == 0) {
            // dt of this step from the current state (cfl_m before the engine)
            CflArgs ca{};
            ca.ro = in[0]; ca.pr = in[1]; ca.vx = in[2]; ca.vy = in[3]; ca.bx = in[4]; ca.by = in[5];
            ca.dx = dx; ca.dy = dy; ca.safety = safety; ca.gm = gm; ca.ix = ix; ca.jx = jx; ca.kx = 1;
            ca.dim = 2; ca.clk = c; ca.step = n;
            HIPCHK(launch_cfl(ca, s));
        }
        Mlw2dArgs a{};
        a.ro = in[0]; a.pr = in[1]; a.vx = in[2]; a.vy = in[3]; a.bx = in[4]; a.by = in[5]; a.az = in[6];
        a.ro_o = out[0]; a.pr_o = out[1]; a.vx_o = out[2]; a.vy_o = out[3]; a.bx_o = out[4];
        a.by_o = out[5]; a.az_o = out[6];
        a.dx = dx; a.dxm = dxm; a.dy = dy; a.dym = dym; a.mx = mx; a.my = my;
        a.qav = qav; a.gm = gm; a.dt = 0.0; a.clk = c; a.step = n; a.ix = ix; a.jx = jx;
        a.per_margin = periodic ? margin : 0;
        a.safety = periodic ? safety : 0.0;
        const int e = n - step0;
        if (events) HIPCHK(hipEventRecord((hipEvent_t)events[2 * e], s));
        HIPCHK(launch_mlw2d(a, s));
        if (events) HIPCHK(hipEventRecord((hipEvent_t)events[2 * e + 1], s));
        if (!periodic) {
            // ghost cells of the new state: bnd (x faces, then y faces)
            for (int dir = 0; dir < 2; dir++) {
                for (int side = 0; side < 2; side++) {
                    BndArgs b{};
                    b.nf = 0;
                    for (int f = 0; f < 7; f++) {
                        const int kind = bc[(dir * 2 + side) * 7 + f];
                        if (kind == BD_NONE) continue;
                        if (kind == BD_PER && side == 1) continue;   // both sides in one pass
                        b.q[b.nf] = out[f]; b.kind[b.nf] = kind; b.q0[b.nf] = 0.0; b.nf++;
                    }
                    if (!b.nf) continue;
                    b.dir = dir; b.margin = margin; b.ix = ix; b.jx = jx; b.kx = 1;
                    b.mbnd = side;
                    b.dm = dir == 0 ? dxm : dym;
                    // periodic fields go with margin0 = margin1 = margin
                    BndArgs bp = b, bo = b;
                    bp.nf = bo.nf = 0;
                    for (int f = 0; f < b.nf; f++) {
                        BndArgs &t = (b.kind[f] == BD_PER) ? bp : bo;
                        t.q[t.nf] = b.q[f]; t.kind[t.nf] = b.kind[f]; t.q0[t.nf] = 0.0; t.nf++;
                    }
                    if (bp.nf) { bp.mbnd = margin; HIPCHK(launch_bnd(bp, s)); }
                    if (bo.nf) HIPCHK(launch_bnd(bo, s));
                }
            }
        }
    }
    return 0;
}

int cans2d_metric_ws_size(int ix, int jx) { return (int)(4 * al(ix) + 4 * al(jx)); }

int cans2d_metric_fill(const double *dx, const double *dxm, int ix, const double *dy, const double *dym, int jx,
                       double *metric_ws, void *stream) {
    if (!dx || !dxm || !dy || !dym || !metric_ws || ix < 3 || jx < 3)
        return fail(CANS_ERR_ARG, "cans2d_metric_fill: bad argument (ix=%d jx=%d)", ix, jx);
    hipStream_t s = (hipStream_t)stream;
    double *m = metric_ws;
    const size_t mxn = al(ix), myn = al(jx);
    Metric mx{m, m + mxn, m + 2 * mxn, m + 3 * mxn};
    m += 4 * mxn;
    Metric my{m, m + myn, m + 2 * myn, m + 3 * myn};
    HIPCHK(launch_metric(dx, dxm, ix, mx, s));
    HIPCHK(launch_metric(dy, dym, jx, my, s));
    return 0;
}

// ====================================================================== adjoint
size_t cans2d_adj_ws_bytes(int ix, int jx) { return adj2d_ws_doubles(ix, jx) * sizeof(double); }

int cans2d_mlw_adjoint_step(const double *const *pn6, const double *const *pn1_6, double *const *lam_in6,
                            double *const *lam_out6, double qav, double gm, double safety, const double *dx,
                            const double *dxm, int ix, const double *dy, const double *dym, int jx, int margin,
                            const double *metric_ws, void *ws, void *stream) {
    if (!pn6 || !pn1_6 || !lam_in6 || !lam_out6 || !dx || !dxm || !dy || !dym || !metric_ws || !ws)
        return fail(CANS_ERR_ARG, "cans2d_mlw_adjoint_step: NULL argument");
    if (ix < 3 || jx < 3 || margin < 1 || ix - 2 * margin < margin || jx - 2 * margin < margin)
        return fail(CANS_ERR_ARG, "cans2d_mlw_adjoint_step: periodic box needs ix-2*margin >= margin (ix=%d jx=%d margin=%d)",
                    ix, jx, margin);
    for (int f = 0; f < 6; f++)
        if (!pn6[f] || !pn1_6[f] || !lam_in6[f] || !lam_out6[f])
            return fail(CANS_ERR_ARG, "cans2d_mlw_adjoint_step: NULL field %d", f);
    hipStream_t s = (hipStream_t)stream;
    const double *m = metric_ws;
    const size_t mxn = al(ix), myn = al(jx);
    Metric mx{(double *)m, (double *)m + mxn, (double *)m + 2 * mxn, (double *)m + 3 * mxn};
    m += 4 * mxn;
    Metric my{(double *)m, (double *)m + myn, (double *)m + 2 * myn, (double *)m + 3 * myn};
    // the step's dt: cfl_m over every cell of P^n, as the forward computed it
    double *tail = (double *)ws + 12 * al((size_t)ix * jx);
    unsigned long long *key = reinterpret_cast<unsigned long long *>(tail);
    HIPCHK(launch_arm_key(key, s));
    CflArgs c{};
    c.ro = pn6[0]; c.pr = pn6[1]; c.vx = pn6[2]; c.vy = pn6[3]; c.bx = pn6[4]; c.by = pn6[5];
    c.dx = dx; c.dy = dy; c.safety = safety; c.gm = gm; c.ix = ix; c.jx = jx; c.kx = 1; c.dim = 2; c.key = key;
    HIPCHK(launch_cfl(c, s));
    HIPCHK(launch_key_to_dt(key, tail + 1, s));
    HIPCHK(launch_zero(tail + 2, kAdjAccSlot, s));
    HIPCHK(launch_adj2d_step(pn6, pn1_6, lam_in6, lam_out6, (double *)ws, tail + 1, tail + 2, nullptr, nullptr,
                             nullptr, dx, dxm, dy, dym, mx, my, qav, gm, safety, ix, jx, margin, s));
    HIPCHK(launch_adj2d_dt(pn6, lam_out6, (double *)ws, tail + 1, tail + 2, dx, dy, gm, safety, ix, jx, s));
    return 0;
}

int cans2d_mlw_record(double *const *states, int nsteps, void *clk, double qav, double gm, double safety,
                      const double *dx, const double *dxm, int ix, const double *dy, const double *dym, int jx,
                      int margin, double *metric_ws, double *dt_log, void *stream) {
    if (!states || !clk || !metric_ws || !dt_log || nsteps < 0)
        return fail(CANS_ERR_ARG, "cans2d_mlw_record: bad argument");
    if (ix < 3 || jx < 3 || margin < 1 || ix - 2 * margin < margin || jx - 2 * margin < margin)
        return fail(CANS_ERR_ARG, "cans2d_mlw_record: periodic box needs ix-2*margin >= margin (ix=%d jx=%d margin=%d)",
                    ix, jx, margin);
    for (int k = 0; k < 7 * (nsteps + 1); k++)
        if (!states[k]) return fail(CANS_ERR_ARG, "cans2d_mlw_record: NULL state pointer %d", k);
    hipStream_t s = (hipStream_t)stream;
    Clock *c = (Clock *)clk;
    double *m = metric_ws;
    const size_t mxn = al(ix), myn = al(jx);
    Metric mx{m, m + mxn, m + 2 * mxn, m + 3 * mxn};
    m += 4 * mxn;
    Metric my{m, m + myn, m + 2 * myn, m + 3 * myn};
    HIPCHK(launch_metric(dx, dxm, ix, mx, s));
    HIPCHK(launch_metric(dy, dym, jx, my, s));
    for (int n = 0; n < nsteps; n++) {
        double *const *in = states + 7 * n;
        double *const *out = states + 7 * (n + 1);
        if (n == 0) {
            CflArgs ca{};
            ca.ro = in[0]; ca.pr = in[1]; ca.vx = in[2]; ca.vy = in[3]; ca.bx = in[4]; ca.by = in[5];
            ca.dx = dx; ca.dy = dy; ca.safety = safety; ca.gm = gm; ca.ix = ix; ca.jx = jx; ca.kx = 1;
            ca.dim = 2; ca.clk = c; ca.step = 0;
            HIPCHK(launch_cfl(ca, s));
        }
        Mlw2dArgs a{};
        a.ro = in[0]; a.pr = in[1]; a.vx = in[2]; a.vy = in[3]; a.bx = in[4]; a.by = in[5]; a.az = in[6];
        a.ro_o = out[0]; a.pr_o = out[1]; a.vx_o = out[2]; a.vy_o = out[3]; a.bx_o = out[4];
        a.by_o = out[5]; a.az_o = out[6];
        a.dx = dx; a.dxm = dxm; a.dy = dy; a.dym = dym; a.mx = mx; a.my = my;
        a.qav = qav; a.gm = gm; a.dt = 0.0; a.clk = c; a.step = n; a.ix = ix; a.jx = jx;
        a.per_margin = margin;
        a.safety = safety;
        a.dt_log = dt_log;
        HIPCHK(launch_mlw2d(a, s));
    }
    return 0;
}

int cans2d_mlw_adjoint_run(const double *const *states, const double *dt_log, int nsteps, double *const *lam6,
                           double *const *tmp6, double qav, double gm, double safety, const double *dx,
                           const double *dxm, int ix, const double *dy, const double *dym, int jx, int margin,
                           const double *metric_ws, void *ws, void *stream) {
    if (!states || !dt_log || !lam6 || !tmp6 || !metric_ws || !ws || nsteps < 0)
        return fail(CANS_ERR_ARG, "cans2d_mlw_adjoint_run: bad argument");
    if (ix < 3 || jx < 3 || margin < 1 || ix - 2 * margin < margin || jx - 2 * margin < margin)
        return fail(CANS_ERR_ARG, "cans2d_mlw_adjoint_run: periodic box needs ix-2*margin >= margin (ix=%d jx=%d margin=%d)",
                    ix, jx, margin);
    for (int k = 0; k < 7 * (nsteps + 1); k++)
        if (!states[k]) return fail(CANS_ERR_ARG, "cans2d_mlw_adjoint_run: NULL state pointer %d", k);
    for (int f = 0; f < 6; f++)
        if (!lam6[f] || !tmp6[f]) return fail(CANS_ERR_ARG, "cans2d_mlw_adjoint_run: NULL lam/tmp field %d", f);
    hipStream_t s = (hipStream_t)stream;
    const double *m = metric_ws;
    const size_t mxn = al(ix), myn = al(jx);
    Metric mx{(double *)m, (double *)m + mxn, (double *)m + 2 * mxn, (double *)m + 3 * mxn};
    m += 4 * mxn;
    Metric my{(double *)m, (double *)m + myn, (double *)m + 2 * myn, (double *)m + 3 * myn};
    double *const *cur = lam6;
    double *const *oth = tmp6;
    double *acc = (double *)ws + 12 * al((size_t)ix * jx) + 4;   // two accumulator slots
    HIPCHK(launch_zero(acc, 2 * kAdjAccSlot, s));
    // step n counts into acc slot n&1; its B1 applies the dt gradient of step
    // n+1 (slot (n+1)&1), which its B2 then clears for step n-1
    for (int n = nsteps - 1; n >= 0; n--) {
        const bool pend = n < nsteps - 1;
        HIPCHK(launch_adj2d_step(states + 7 * n, states + 7 * (n + 1), cur, oth, (double *)ws, dt_log + n,
                                 acc + kAdjAccSlot * (n & 1), pend ? dt_log + n + 1 : nullptr,
                                 pend ? acc + kAdjAccSlot * ((n + 1) & 1) : nullptr,
                                 acc + kAdjAccSlot * ((n + 1) & 1), dx, dxm, dy,
                                 dym, mx, my, qav, gm, safety, ix, jx, margin, s));
        double *const *t = cur;
        cur = oth;
        oth = t;
    }
    if (nsteps > 0)   // the dt gradient of the segment's first step
        HIPCHK(launch_adj2d_dt(states, cur, (double *)ws, dt_log, acc, dx, dy, gm, safety, ix, jx, s));
    if (cur != lam6) HIPCHK(launch_copy6(lam6, cur, (size_t)ix * jx, s));
    return 0;
}

size_t cans2d_roe_adj_ws_bytes(int ix, int jx) { return adj2d_roe_ws_doubles(ix, jx) * sizeof(double); }

int cans2d_roe_adjoint_run(const double *const *states, const double *const *half, const double *dt_log, int nsteps,
                           double *const *lam6, double *const *tmp6, double gm, double safety, const double *dx,
                           const double *dy, int ix, int jx, int margin, void *ws, void *stream) {
    if (!states || !dt_log || !lam6 || !tmp6 || !dx || !dy || !ws || nsteps < 0)
        return fail(CANS_ERR_ARG, "cans2d_roe_adjoint_run: bad argument");
    if (ix < 8 || jx < 8 || margin < 3 || ix - 2 * margin < margin || jx - 2 * margin < margin)
        return fail(CANS_ERR_ARG, "cans2d_roe_adjoint_run: the Roe box needs margin >= 3 and ix-2*margin >= margin "
                    "(ix=%d jx=%d margin=%d)", ix, jx, margin);
    for (int k = 0; k < 7 * (nsteps + 1); k++)
        if (!states[k]) return fail(CANS_ERR_ARG, "cans2d_roe_adjoint_run: NULL state pointer %d", k);
    for (int f = 0; f < 6; f++)
        if (!lam6[f] || !tmp6[f]) return fail(CANS_ERR_ARG, "cans2d_roe_adjoint_run: NULL lam/tmp field %d", f);
    hipStream_t s = (hipStream_t)stream;
    double *const *cur = lam6;
    double *const *oth = tmp6;
    const size_t n = (size_t)ix * jx;
    double *acc = (double *)ws + adj2d_roe_ws_doubles(ix, jx) - 64 + 4;   // two accumulator slots
    HIPCHK(launch_zero(acc, 2 * kAdjAccSlot, s));
    // as cans2d_mlw_adjoint_run: step n counts into slot n&1, its B1 applies
    // the pending dt gradient of step n+1, which its R2 then clears
    for (int st = nsteps - 1; st >= 0; st--) {
        const bool pend = st < nsteps - 1;
        HIPCHK(launch_adj2d_roe_step(states + 7 * st, states + 7 * (st + 1), half ? half + 6 * st : nullptr, cur, oth,
                                     (double *)ws, dt_log + st,
                                     acc + kAdjAccSlot * (st & 1), pend ? dt_log + st + 1 : nullptr,
                                     pend ? acc + kAdjAccSlot * ((st + 1) & 1) : nullptr,
                                     acc + kAdjAccSlot * ((st + 1) & 1), dx, dy, gm, safety, ix, jx, margin, s));
        double *const *t = cur;
        cur = oth;
        oth = t;
    }
    if (nsteps > 0)   // the dt gradient of the first step
        HIPCHK(launch_adj2d_dt(states, cur, (double *)ws, dt_log, acc, dx, dy, gm, safety, ix, jx, s));
    if (cur != lam6) HIPCHK(launch_copy6(lam6, cur, n, s));
    return 0;
}

int cans2d_periodic_fold(double *const *lam6, int ix, int jx, int margin, void *stream) {
    if (!lam6 || ix < 3 || jx < 3 || margin < 1 || ix - 2 * margin < margin || jx - 2 * margin < margin)
        return fail(CANS_ERR_ARG, "cans2d_periodic_fold: bad argument (ix=%d jx=%d margin=%d)", ix, jx, margin);
    HIPCHK(launch_periodic_fold(lam6, ix, jx, margin, (hipStream_t)stream));
    return 0;
}

// ====================================================================== events
int cans_event_create(void **ev) {
    HIPCHK(hipEventCreate((hipEvent_t *)ev));
    return 0;
}
int cans_event_destroy(void *ev) {
    if (ev) HIPCHK(hipEventDestroy((hipEvent_t)ev));
    return 0;
}
int cans_event_record(void *ev, void *stream) {
    HIPCHK(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
    return 0;
}
int cans_event_elapsed_ms(void *start, void *stop, float *ms) {
    HIPCHK(hipEventSynchronize((hipEvent_t)stop));
    HIPCHK(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return 0;
}

// ====================================================================== host mapping
// Device mirrors of host arrays for the Fortran drop-in wrappers.
namespace {
struct Mirror {
    void *dev = nullptr;
    size_t cap = 0;       // device allocation
    size_t bytes = 0;     // size of the current mapping (what unmap copies back)
    bool valid = false;   // device copy is current
    bool dirty = false;   // resident mode: a routine changed the device copy since the last sync
};
std::map<const void *, Mirror> g_mirror;
int g_resident = 0;
}  // namespace

void *cans_gpu_map(const void *host, size_t bytes, int copy_in) {
    std::lock_guard<std::mutex> lk(g_mu);
    Mirror &m = g_mirror[host];
    if (m.cap < bytes) {
        if (m.dev) (void)hipFree(m.dev);
        m.dev = nullptr;
        if (hipMalloc(&m.dev, bytes) != hipSuccess) {
            fail(CANS_ERR_HIP, "cans_gpu_map: hipMalloc(%zu) failed", bytes);
            g_mirror.erase(host);
            return nullptr;
        }
        m.cap = bytes;
        m.valid = false;
    }
    // A host address can come back for an array of another size (the
    // allocator reuses it): the mapping's size is this call's, so that unmap
    // never copies past the end of the host array, and a size change makes
    // the device copy stale.
    if (m.bytes != bytes) {
        m.valid = false;
        m.dirty = false;   // the device data belonged to the array that lived here before
    }
    m.bytes = bytes;
    const bool need = copy_in && (!g_resident || !m.valid);
    if (need) {
        if (hipMemcpy(m.dev, host, bytes, hipMemcpyHostToDevice) != hipSuccess) {
            fail(CANS_ERR_HIP, "cans_gpu_map: H2D copy failed");
            return nullptr;
        }
    }
    if (copy_in) m.valid = true;
    return m.dev;
}

int cans_gpu_unmap(void *host, int copy_out) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_mirror.find(host);
    if (it == g_mirror.end()) return fail(CANS_ERR_ARG, "cans_gpu_unmap: %p not mapped", host);
    if (copy_out && !g_resident) HIPCHK(hipMemcpy(host, it->second.dev, it->second.bytes, hipMemcpyDeviceToHost));
    if (copy_out && g_resident) it->second.dirty = true;
    if (!copy_out && !g_resident) HIPCHK(hipDeviceSynchronize());
    return 0;
}

// Copy back the mirrors a routine changed since the last sync (and only
// those: read-only inputs and arrays the program has since released stay
// untouched on the host).  Caller holds g_mu.
static int flush_dirty() {
    HIPCHK(hipDeviceSynchronize());
    for (auto &kv : g_mirror) {
        Mirror &m = kv.second;
        if (!m.valid || !m.dirty) continue;
        HIPCHK(hipMemcpy(const_cast<void *>(kv.first), m.dev, m.bytes, hipMemcpyDeviceToHost));
        m.dirty = false;
    }
    return 0;
}

int cans_gpu_set_resident(int on) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_resident && !on)
        if (int rc = flush_dirty()) return rc;   // leaving resident mode: the host gets the results first
    g_resident = on ? 1 : 0;
    if (!on)
        for (auto &kv : g_mirror) kv.second.valid = false;
    return 0;
}

int cans_gpu_sync_host(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    return flush_dirty();
}

int cans_gpu_invalidate(const void *host) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_mirror.find(host);
    if (it != g_mirror.end()) {
        it->second.valid = false;
        it->second.dirty = false;
    }
    return 0;
}

int cans_gpu_forget(const void *host) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_mirror.find(host);
    if (it == g_mirror.end()) return 0;
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipFree(it->second.dev));
    g_mirror.erase(it);
    return 0;
}

int cans_gpu_release_all(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    (void)hipDeviceSynchronize();
    for (auto &kv : g_mirror) (void)hipFree(kv.second.dev);
    g_mirror.clear();
    for (auto &kv : g_pool) (void)hipFree(kv.second.ptr);
    g_pool.clear();
    return 0;
}

}  // extern "C"
