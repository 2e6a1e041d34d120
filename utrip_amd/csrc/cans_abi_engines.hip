// cans_abi_engines.hip -- extern "C" entry points of the Roe-MUSCL, 3D MLW
// and HLLD engines, and the resident multi-step runs of every engine
// (declared in include/cans_gpu.h).
#include <hip/hip_runtime.h>
#include <string.h>

#include "cans_common.hpp"
#include "cans_kernels.hpp"
#include "cans_roe.hpp"
#include "../../include/cans_gpu.h"

namespace cans {
int abi_fail(int code, const char *fmt, ...);
int abi_scratch(size_t bytes, void **out, hipStream_t s);
int host_trig(const double *y, const double *ym, int jx, double *dev, hipStream_t s);
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

inline size_t al(size_t n) { return (n + 63) & ~size_t(63); }

// carve `count` arrays of n doubles out of a workspace cursor
struct Carver {
    double *p;
    double *take(size_t n) {
        double *r = p;
        p += al(n);
        return r;
    }
};

size_t metric_doubles(int ix, int jx, int kx) { return 4 * (al(ix) + al(jx) + al(kx)); }

void carve_metric(Carver &c, int ix, int jx, int kx, Metric *m) {
    const int n[3] = {ix, jx, kx};
    for (int d = 0; d < 3; d++) {
        m[d].di = c.take(n[d]);
        m[d].dim = c.take(n[d]);
        m[d].u0 = c.take(n[d]);
        m[d].u1 = c.take(n[d]);
    }
}

// engine scratch (doubles) beyond the metric
size_t engine_doubles(int dim, int engine, size_t N) {
    const size_t a = al(N);
    switch (engine) {
    case CANS_ENGINE_MLW: return dim == 3 ? (size_t)(kMlw3dCellFields + kMlw3dCornerFields) * a : 0;
    case CANS_ENGINE_MLWC: return dim == 3 ? (size_t)(kMlw3dCylCellFields + kMlw3dCylCornerFields) * a : 0;
    case CANS_ENGINE_MLWS:
        return (size_t)(dim == 3 ? kMlw3dCylCellFields + kMlw3dCylCornerFields : kMlw2sCellFields + kMlw2sCornerFields) * a;
    case CANS_ENGINE_ROE:
    case CANS_ENGINE_ROE3: return (8 + 3 * 8) * a;
    case CANS_ENGINE_HLLD: return 3 * 8 * a;
    default: return 0;
    }
}

void roe_scratch(RoeArgs &r, Carver &c, size_t N, int dim, bool comp3 = false) {
    const bool z = dim == 3 || comp3;
    const bool used[8] = {true, true, true, true, z, dim >= 2, true, z};
    for (int q = 0; q < 8; q++) r.half[q] = used[q] ? c.take(N) : (c.take(N), nullptr);
    for (int d = 0; d < 3; d++)
        for (int q = 0; q < 8; q++) r.flux[d][q] = c.take(N);
}

void hlld_scratch(HlldArgs &h, Carver &c, size_t N) {
    for (int q = 0; q < 8; q++) h.half[q] = c.take(N);
    for (int q = 0; q < 8; q++) h.u1[q] = c.take(N);
    for (int q = 0; q < 8; q++) h.u2[q] = c.take(N);
}

int check_margin(const char *who, int n, int lo) {
    if (n < lo) return abi_fail(CANS_ERR_ARG, "%s: dimension %d < %d", who, n, lo);
    return 0;
}

}  // namespace

extern "C" {

// ====================================================================== Roe drop-in
static int roe_dropin(int dim, bool mixed, double *const *prim, double *az, const double *bx, const double *bxm,
                      double dt, double gm, const double *dx, int ix, const double *dy, int jx, const double *dz,
                      int kx, void *stream, bool comp3 = false) {
    const int n[3] = {ix, jx, kx};
    for (int d = 0; d < dim; d++)
        if (int rc = check_margin("roe_m", n[d], 5)) return rc;
    hipStream_t s = (hipStream_t)stream;
    const size_t N = (size_t)ix * jx * kx;
    void *w = nullptr;
    if (int rc = abi_scratch((engine_doubles(dim, CANS_ENGINE_ROE, N) + 9 * al(N)) * sizeof(double), &w, s)) return rc;
    Carver c{(double *)w};
    RoeArgs r{};
    for (int q = 0; q < 8; q++) r.prim[q] = prim[q];
    roe_scratch(r, c, N, dim, comp3);
    double *outbuf[8];
    for (int q = 0; q < 8; q++) { outbuf[q] = c.take(N); r.out[q] = prim[q] ? outbuf[q] : nullptr; }
    double *azo = c.take(N);
    r.az = az; r.out_az = dim == 2 ? azo : nullptr;
    r.bx = bx; r.bxm = bxm; r.dx = dx; r.dy = dy; r.dz = dz;
    r.gm = gm; r.dt = dt; r.clk = nullptr; r.ix = ix; r.jx = jx; r.kx = kx;
    HIPCHK(launch_roe(r, dim, mixed, s, comp3));
    double *dst[9];
    const double *src[9];
    int nf = 0;
    for (int q = 0; q < 8; q++)
        if (prim[q]) { dst[nf] = prim[q]; src[nf] = outbuf[q]; nf++; }
    if (dim == 2) { dst[nf] = az; src[nf] = azo; nf++; }
    HIPCHK(launch_copy_interior(dst, src, nf, ix, jx, kx, 3, s));
    return 0;
}

int cans1d_roe_m(double *ro, double *pr, double *vx, double *vy, double *by, const double *bx,
                 const double *bxm, double dt, double gm, const double *dx, int ix, void *stream) {
    double *prim[8] = {ro, pr, vx, vy, nullptr, nullptr, by, nullptr};
    return roe_dropin(1, false, prim, nullptr, bx, bxm, dt, gm, dx, ix, nullptr, 1, nullptr, 1, stream);
}

int cans2d_roe_m(double *ro, double *pr, double *vx, double *vy, double *bx, double *by, double *az,
                 double dt, double gm, const double *dx, int ix, const double *dy, int jx, void *stream) {
    double *prim[8] = {ro, pr, vx, vy, nullptr, bx, by, nullptr};
    return roe_dropin(2, false, prim, az, nullptr, nullptr, dt, gm, dx, ix, dy, jx, nullptr, 1, stream);
}

int cans1d_roe_m3(double *ro, double *pr, double *vx, double *vy, double *vz, double *by, double *bz,
                  const double *bx, const double *bxm, double dt, double gm, const double *dx, int ix,
                  void *stream) {
    double *prim[8] = {ro, pr, vx, vy, vz, nullptr, by, bz};
    return roe_dropin(1, false, prim, nullptr, bx, bxm, dt, gm, dx, ix, nullptr, 1, nullptr, 1, stream, true);
}

int cans2d_roe_m3(double *ro, double *pr, double *vx, double *vy, double *vz, double *bx, double *by, double *bz,
                  double *az, double dt, double gm, const double *dx, int ix, const double *dy, int jx,
                  void *stream) {
    double *prim[8] = {ro, pr, vx, vy, vz, bx, by, bz};
    return roe_dropin(2, false, prim, az, nullptr, nullptr, dt, gm, dx, ix, dy, jx, nullptr, 1, stream, true);
}

int cans3d_roe_m(double *ro, double *pr, double *vx, double *vy, double *vz, double *bx, double *by,
                 double *bz, double dt, double gm, const double *dx, int ix, const double *dy, int jx,
                 const double *dz, int kx, void *stream) {
    double *prim[8] = {ro, pr, vx, vy, vz, bx, by, bz};
    return roe_dropin(3, false, prim, nullptr, nullptr, nullptr, dt, gm, dx, ix, dy, jx, dz, kx, stream);
}

int cans3d_roe_m_mixed(double *ro, double *pr, double *vx, double *vy, double *vz, double *bx, double *by,
                       double *bz, double dt, double gm, const double *dx, int ix, const double *dy, int jx,
                       const double *dz, int kx, void *stream) {
    double *prim[8] = {ro, pr, vx, vy, vz, bx, by, bz};
    return roe_dropin(3, true, prim, nullptr, nullptr, nullptr, dt, gm, dx, ix, dy, jx, dz, kx, stream);
}

// ====================================================================== 3D MLW drop-in
// mlw_m (x = NULL), the cylindrical mlw_m_c (cell/face radii x, xm) or the
// spherical mlw_m_s (and the colatitudes y, ym)
static int mlw3d_dropin(const char *who, double *const *prim, const double *x, const double *xm, const double *y,
                        const double *ym, double dt, double qav, double gm, const double *dx, const double *dxm,
                        int ix, const double *dy, const double *dym, int jx, const double *dz, const double *dzm,
                        int kx, void *stream) {
    if (ix < 3 || jx < 3 || kx < 3) return abi_fail(CANS_ERR_ARG, "%s: %d x %d x %d", who, ix, jx, kx);
    for (int q = 0; q < 8; q++)
        if (!prim[q]) return abi_fail(CANS_ERR_ARG, "%s: NULL field %d", who, q);
    const bool cyl = x != nullptr, sph = y != nullptr;
    if (cyl && !xm) return abi_fail(CANS_ERR_ARG, "%s: NULL xm", who);
    if (sph && (!ym || !cyl)) return abi_fail(CANS_ERR_ARG, "%s: NULL ym or radii", who);
    hipStream_t s = (hipStream_t)stream;
    const size_t N = (size_t)ix * jx * kx;
    void *w = nullptr;
    const int engine = cyl ? CANS_ENGINE_MLWC : CANS_ENGINE_MLW;
    const size_t need =
        metric_doubles(ix, jx, kx) + engine_doubles(3, engine, N) + 8 * al(N) + (sph ? 6 * al(jx) : 0);
    if (int rc = abi_scratch(need * sizeof(double), &w, s)) return rc;
    Carver c{(double *)w};
    Metric m[3];
    carve_metric(c, ix, jx, kx, m);
    HIPCHK(launch_metric(dx, dxm, ix, m[0], s));
    HIPCHK(launch_metric(dy, dym, jx, m[1], s));
    HIPCHK(launch_metric(dz, dzm, kx, m[2], s));
    Mlw3dArgs a{};
    for (int q = 0; q < 8; q++) a.prim[q] = prim[q];
    a.cell = c.take((size_t)(cyl ? kMlw3dCylCellFields : kMlw3dCellFields) * al(N));
    a.corner = c.take((size_t)(cyl ? kMlw3dCylCornerFields : kMlw3dCornerFields) * al(N));
    for (int q = 0; q < 8; q++) a.out[q] = c.take(N);
    a.x = x; a.xm = xm;
    if (sph) {
        double *tab = c.take(6 * (size_t)jx);
        if (int rc = host_trig(y, ym, jx, tab, s)) return rc;
        a.trig = tab;
    }
    a.dxm = dxm; a.dym = dym; a.dzm = dzm; a.mx = m[0]; a.my = m[1]; a.mz = m[2];
    a.qav = qav; a.gm = gm; a.dt = dt; a.clk = nullptr; a.ix = ix; a.jx = jx; a.kx = kx;
    HIPCHK(launch_mlw3d(a, s));
    const double *src[8];
    for (int q = 0; q < 8; q++) src[q] = a.out[q];
    HIPCHK(launch_copy_interior(prim, src, 8, ix, jx, kx, 2, s));
    return 0;
}

int cans3d_mlw_m(double *ro, double *pr, double *vx, double *vy, double *vz, double *bx, double *by,
                 double *bz, double dt, double qav, double gm, const double *dx, const double *dxm, int ix,
                 const double *dy, const double *dym, int jx, const double *dz, const double *dzm, int kx,
                 void *stream) {
    double *prim[8] = {ro, pr, vx, vy, vz, bx, by, bz};
    return mlw3d_dropin("cans3d_mlw_m", prim, nullptr, nullptr, nullptr, nullptr, dt, qav, gm, dx, dxm, ix, dy, dym,
                        jx, dz, dzm, kx, stream);
}

int cans3d_mlw_m_c(double *ro, double *pr, double *vx, double *vy, double *vz, double *bx, double *by,
                   double *bz, double dt, double qav, double gm, const double *x, const double *xm,
                   const double *dx, const double *dxm, int ix, const double *dy, const double *dym, int jx,
                   const double *dz, const double *dzm, int kx, void *stream) {
    if (!x || !xm) return abi_fail(CANS_ERR_ARG, "cans3d_mlw_m_c: NULL radius");
    double *prim[8] = {ro, pr, vx, vy, vz, bx, by, bz};
    return mlw3d_dropin("cans3d_mlw_m_c", prim, x, xm, nullptr, nullptr, dt, qav, gm, dx, dxm, ix, dy, dym, jx, dz,
                        dzm, kx, stream);
}

int cans3d_mlw_m_s(double *ro, double *pr, double *vx, double *vy, double *vz, double *bx, double *by,
                   double *bz, double dt, double qav, double gm, const double *x, const double *xm,
                   const double *dx, const double *dxm, int ix, const double *y, const double *ym,
                   const double *dy, const double *dym, int jx, const double *dz, const double *dzm, int kx,
                   void *stream) {
    if (!x || !xm || !y || !ym) return abi_fail(CANS_ERR_ARG, "cans3d_mlw_m_s: NULL geometry array");
    double *prim[8] = {ro, pr, vx, vy, vz, bx, by, bz};
    return mlw3d_dropin("cans3d_mlw_m_s", prim, x, xm, y, ym, dt, qav, gm, dx, dxm, ix, dy, dym, jx, dz, dzm, kx,
                        stream);
}

// ====================================================================== mlw_m3 drop-in
static int mlw3_dropin(int dim, double *const *f, const double *bx, const double *bxm, double dt, double qav,
                       double gm, const double *dx, const double *dxm, int ix, const double *dy, const double *dym,
                       int jx, void *stream) {
    const int nf = dim == 1 ? 7 : 9;
    if (ix < 3 || (dim == 2 && jx < 3)) return abi_fail(CANS_ERR_ARG, "mlw_m3: %d x %d", ix, jx);
    for (int q = 0; q < nf; q++)
        if (!f[q]) return abi_fail(CANS_ERR_ARG, "mlw_m3: NULL field %d", q);
    if (dim == 1 && (!bx || !bxm)) return abi_fail(CANS_ERR_ARG, "mlw_m3: NULL bx/bxm");
    hipStream_t s = (hipStream_t)stream;
    const size_t N = (size_t)ix * (dim == 2 ? jx : 1);
    void *w = nullptr;
    if (int rc = abi_scratch((metric_doubles(ix, jx, 1) + nf * al(N)) * sizeof(double), &w, s)) return rc;
    Carver c{(double *)w};
    Metric m[3];
    carve_metric(c, ix, jx, 1, m);
    HIPCHK(launch_metric(dx, dxm, ix, m[0], s));
    if (dim == 2) HIPCHK(launch_metric(dy, dym, jx, m[1], s));
    Mlw3Args a{};
    for (int q = 0; q < nf; q++) { a.in[q] = f[q]; a.out[q] = c.take(N); }
    a.bx = bx; a.bxm = bxm; a.dxm = dxm; a.dym = dym; a.mx = m[0]; a.my = m[1];
    a.qav = qav; a.gm = gm; a.dt = dt; a.clk = nullptr; a.ix = ix; a.jx = dim == 2 ? jx : 1;
    HIPCHK(launch_mlw_m3(a, dim, s));
    const double *src[9];
    for (int q = 0; q < nf; q++) src[q] = a.out[q];
    HIPCHK(launch_copy_interior(f, src, nf, ix, dim == 2 ? jx : 1, 1, 2, s));
    return 0;
}

int cans2d_mlw_m_c(double *ro, double *pr, double *vx, double *vz, double *bx, double *bz, double *ay, double dt,
                   double qav, double gm, const double *x, const double *xm, const double *dx, const double *dxm,
                   int ix, const double *dz, const double *dzm, int jx, void *stream) {
    double *f[7] = {ro, pr, vx, vz, bx, bz, ay};
    for (int q = 0; q < 7; q++)
        if (!f[q]) return abi_fail(CANS_ERR_ARG, "mlw_m_c: NULL field %d", q);
    if (!x || !xm || !dx || !dxm || !dz || !dzm) return abi_fail(CANS_ERR_ARG, "mlw_m_c: NULL grid array");
    if (ix < 3 || jx < 3) return abi_fail(CANS_ERR_ARG, "mlw_m_c: %d x %d", ix, jx);
    hipStream_t s = (hipStream_t)stream;
    const size_t N = (size_t)ix * jx;
    void *w = nullptr;
    if (int rc = abi_scratch((metric_doubles(ix, jx, 1) + 7 * al(N)) * sizeof(double), &w, s)) return rc;
    Carver c{(double *)w};
    Metric m[3];
    carve_metric(c, ix, jx, 1, m);
    HIPCHK(launch_metric(dx, dxm, ix, m[0], s));
    HIPCHK(launch_metric(dz, dzm, jx, m[1], s));
    Mlw3Args a{};
    for (int q = 0; q < 7; q++) { a.in[q] = f[q]; a.out[q] = c.take(N); }
    a.x = x; a.xm = xm; a.dxm = dxm; a.dym = dzm; a.mx = m[0]; a.my = m[1];
    a.qav = qav; a.gm = gm; a.dt = dt; a.clk = nullptr; a.ix = ix; a.jx = jx;
    HIPCHK(launch_mlw_m_c(a, s));
    const double *src[7];
    for (int q = 0; q < 7; q++) src[q] = a.out[q];
    HIPCHK(launch_copy_interior(f, src, 7, ix, jx, 1, 2, s));
    return 0;
}

int cans2d_mlw_m_s(double *ro, double *pr, double *vx, double *vy, double *bx, double *by, double *az, double dt,
                   double qav, double gm, const double *x, const double *xm, const double *dx, const double *dxm,
                   int ix, const double *y, const double *ym, const double *dy, const double *dym, int jx,
                   void *stream) {
    double *f[7] = {ro, pr, vx, vy, bx, by, az};
    for (int q = 0; q < 7; q++)
        if (!f[q]) return abi_fail(CANS_ERR_ARG, "mlw_m_s: NULL field %d", q);
    if (!x || !xm || !dx || !dxm || !y || !ym || !dy || !dym) return abi_fail(CANS_ERR_ARG, "mlw_m_s: NULL grid array");
    if (ix < 3 || jx < 3) return abi_fail(CANS_ERR_ARG, "mlw_m_s: %d x %d", ix, jx);
    hipStream_t s = (hipStream_t)stream;
    const size_t N = (size_t)ix * jx;
    void *w = nullptr;
    const size_t need = metric_doubles(ix, jx, 1) + engine_doubles(2, CANS_ENGINE_MLWS, N) + 7 * al(N) + 6 * al(jx);
    if (int rc = abi_scratch(need * sizeof(double), &w, s)) return rc;
    Carver c{(double *)w};
    Metric m[3];
    carve_metric(c, ix, jx, 1, m);
    HIPCHK(launch_metric(dx, dxm, ix, m[0], s));
    HIPCHK(launch_metric(dy, dym, jx, m[1], s));
    Mlw2sArgs a{};
    for (int q = 0; q < 7; q++) { a.prim[q] = f[q]; a.out[q] = c.take(N); }
    a.cell = c.take((size_t)kMlw2sCellFields * al(N));
    a.corner = c.take((size_t)kMlw2sCornerFields * al(N));
    double *tab = c.take(6 * (size_t)jx);
    if (int rc = host_trig(y, ym, jx, tab, s)) return rc;
    a.x = x; a.xm = xm; a.trig = tab; a.dxm = dxm; a.dym = dym; a.mx = m[0]; a.my = m[1];
    a.qav = qav; a.gm = gm; a.dt = dt; a.clk = nullptr; a.ix = ix; a.jx = jx;
    HIPCHK(launch_mlw2d_s(a, s));
    const double *src[7];
    for (int q = 0; q < 7; q++) src[q] = a.out[q];
    HIPCHK(launch_copy_interior(f, src, 7, ix, jx, 1, 2, s));
    return 0;
}

int cans1d_mlw_m3(double *ro, double *pr, double *vx, double *vy, double *vz, double *by, double *bz,
                  const double *bx, const double *bxm, double dt, double qav, double gm, const double *dx,
                  const double *dxm, int ix, void *stream) {
    double *f[7] = {ro, pr, vx, vy, vz, by, bz};
    return mlw3_dropin(1, f, bx, bxm, dt, qav, gm, dx, dxm, ix, nullptr, nullptr, 1, stream);
}

int cans2d_mlw_m3(double *ro, double *pr, double *vx, double *vy, double *vz, double *bx, double *by, double *bz,
                  double *az, double dt, double qav, double gm, const double *dx, const double *dxm, int ix,
                  const double *dy, const double *dym, int jx, void *stream) {
    double *f[9] = {ro, pr, vx, vy, vz, bx, by, bz, az};
    return mlw3_dropin(2, f, nullptr, nullptr, dt, qav, gm, dx, dxm, ix, dy, dym, jx, stream);
}

// ====================================================================== HLLD drop-in
static int hlld_dropin(bool mixed, double *const *prim, double gm, double mu, double dt, const double *dx,
                       const double *dy, const double *dz, int ix, int jx, int kx, const int *mfdim, void *stream) {
    if (!mfdim) return abi_fail(CANS_ERR_ARG, "hlld: mfdim is NULL");
    const int n[3] = {ix, jx, kx};
    for (int d = 0; d < 3; d++)
        if (mfdim[d] == 1 && n[d] < 5) return abi_fail(CANS_ERR_ARG, "hlld: direction %d has %d cells", d, n[d]);
    hipStream_t s = (hipStream_t)stream;
    const size_t N = (size_t)ix * jx * kx;
    void *w = nullptr;
    if (int rc = abi_scratch(engine_doubles(3, CANS_ENGINE_HLLD, N) * sizeof(double), &w, s)) return rc;
    Carver c{(double *)w};
    HlldArgs h{};
    for (int q = 0; q < 8; q++) { h.prim[q] = prim[q]; h.out[q] = prim[q]; }
    hlld_scratch(h, c, N);
    h.dx = dx; h.dy = dy; h.dz = dz; h.gm = gm; h.mu = mu; h.dt = dt;
    for (int d = 0; d < 3; d++) h.mfdim[d] = mfdim[d];
    h.clk = nullptr; h.ix = ix; h.jx = jx; h.kx = kx;
    HIPCHK(launch_hlld(h, mixed, s));
    return 0;
}

int cans3d_hlld_engine(double *ro, double *pr, double *vx, double *vy, double *vz, double *bx, double *by,
                       double *bz, double gm, double mu, double dt, const double *dx, const double *dy,
                       const double *dz, int ix, int jx, int kx, const int *mfdim, void *stream) {
    double *prim[8] = {ro, pr, vx, vy, vz, bx, by, bz};
    return hlld_dropin(false, prim, gm, mu, dt, dx, dy, dz, ix, jx, kx, mfdim, stream);
}

int cans3d_hlld_engine_mixed(double *ro, double *pr, double *vx, double *vy, double *vz, double *bx, double *by,
                             double *bz, double gm, double mu, double dt, const double *dx, const double *dy,
                             const double *dz, int ix, int jx, int kx, const int *mfdim, void *stream) {
    double *prim[8] = {ro, pr, vx, vy, vz, bx, by, bz};
    return hlld_dropin(true, prim, gm, mu, dt, dx, dy, dz, ix, jx, kx, mfdim, stream);
}

// ====================================================================== resident runs
size_t cans_run_ws_bytes(int dim, int engine, int ix, int jx, int kx) {
    const size_t N = (size_t)ix * jx * kx;
    const size_t trig = engine == CANS_ENGINE_MLWS ? 6 * al(jx) : 0;
    return (metric_doubles(ix, jx, kx) + trig + engine_doubles(dim, engine, N) + 64) * sizeof(double);
}

// One resident run of `nsteps` CANS loop iterations {cfl_m; t=t+dt; engine;
// bnd} for any dimension/engine with generic boundaries.  State set a holds
// step n (n even), b step n+1, alternating.  nf fields per set:
// 1D 5 (ro pr vx vy by), 2D 7 (ro pr vx vy bx by az), 3D 8 (ro pr vx vy vz bx by bz);
// mlw_m3 and roe_m3: 1D 7 (ro pr vx vy vz by bz), 2D 9 (ro pr vx vy vz bx by bz az);
// mlw_m_c: 2D 7 (ro pr vx vz bx bz ay) on the (x, z) plane (dy, dym are dz,
// dzm; bx1d, bxm1d carry the radii x, xm), cfl_m (2D) on (vx, vz, bx, bz);
// 3D 8 (ro pr vx vy vz bx by bz) on (r, phi, z), radii in bx1d, bxm1d,
// cfl_m (3D) as the md_mhdsn 3D driver calls it;
// mlw_m_s: 2D 7 (ro pr vx vy bx by az) on (r, theta) or 3D 8 on (r, theta,
// phi), bx1d = x, xm, y, ym back to back, with cfl_m_s.
int cans2d_roe_record(double *const *states, double *const *half, int nsteps, void *clk, double gm, double safety,
                      const double *dx, const double *dy, int ix, int jx, int margin, void *ws, double *dt_log,
                      void *stream) {
    if (!states || !half || !clk || !ws || !dt_log || !dx || !dy || nsteps < 0)
        return abi_fail(CANS_ERR_ARG, "cans2d_roe_record: NULL argument");
    if (ix < 8 || jx < 8 || margin < 3 || ix - 2 * margin < margin || jx - 2 * margin < margin)
        return abi_fail(CANS_ERR_ARG, "cans2d_roe_record: the Roe box needs margin >= 3 and n-2*margin >= margin "
                        "(ix=%d jx=%d margin=%d)", ix, jx, margin);
    for (int k = 0; k < 7 * (nsteps + 1); k++)
        if (!states[k]) return abi_fail(CANS_ERR_ARG, "cans2d_roe_record: NULL state pointer %d", k);
    for (int k = 0; k < 6 * nsteps; k++)
        if (!half[k]) return abi_fail(CANS_ERR_ARG, "cans2d_roe_record: NULL half-step pointer %d", k);
    hipStream_t s = (hipStream_t)stream;
    Clock *c = (Clock *)clk;
    const size_t N = (size_t)ix * jx;
    Carver cv{(double *)ws};
    Metric m[3];
    carve_metric(cv, ix, jx, 1, m);
    const int map[6] = {0, 1, 2, 3, 5, 6};   // ro pr vx vy bx by -> engine slots
    for (int st = 0; st < nsteps; st++) {
        double *const *in = states + 7 * st;
        double *const *out = states + 7 * (st + 1);
        CflArgs ca{};
        ca.ro = in[0]; ca.pr = in[1]; ca.vx = in[2]; ca.vy = in[3]; ca.bx = in[4]; ca.by = in[5];
        ca.dx = dx; ca.dy = dy; ca.safety = safety; ca.gm = gm;
        ca.ix = ix; ca.jx = jx; ca.kx = 1; ca.dim = 2; ca.clk = c; ca.step = st;
        HIPCHK(launch_cfl(ca, s));
        HIPCHK(launch_clock_step(c, st, s));
        HIPCHK(launch_clock_dt_log(c, st, dt_log + st, s));
        RoeArgs r{};
        for (int f = 0; f < 6; f++) { r.prim[map[f]] = in[f]; r.out[map[f]] = out[f]; }
        Carver ec{cv.p};
        roe_scratch(r, ec, N, 2);
        for (int f = 0; f < 6; f++) r.half[map[f]] = half[6 * st + f];   // the half step goes to the record
        r.az = in[6];
        r.out_az = out[6];
        r.dx = dx; r.dy = dy; r.gm = gm; r.clk = c; r.step = st; r.ix = ix; r.jx = jx; r.kx = 1;
        HIPCHK(launch_roe(r, 2, false, s));
        for (int dir = 0; dir < 2; dir++) {
            BndArgs bp{};
            for (int f = 0; f < 7; f++) { bp.q[bp.nf] = out[f]; bp.kind[bp.nf] = BD_PER; bp.nf++; }
            bp.dir = dir; bp.mbnd = margin; bp.margin = margin; bp.dm = dir == 0 ? dx : dy;
            bp.ix = ix; bp.jx = jx; bp.kx = 1;
            HIPCHK(launch_bnd(bp, s));
        }
    }
    return 0;
}

int cans_run(int dim, int engine, int mixed, double *const *a, double *const *b, const double *bx1d,
             const double *bxm1d, void *clk, int step0, int nsteps, double qav, double gm, double mu,
             double safety, const double *dx, const double *dxm, int ix, const double *dy, const double *dym,
             int jx, const double *dz, const double *dzm, int kx, const int *bc, int margin, void *ws,
             void *const *events, void *stream) {
    if (dim < 1 || dim > 3) return abi_fail(CANS_ERR_ARG, "cans_run: dim %d", dim);
    if (!a || !b || !clk || !ws) return abi_fail(CANS_ERR_ARG, "cans_run: NULL argument");
    if (engine == CANS_ENGINE_HLLD && dim != 3) return abi_fail(CANS_ERR_ARG, "cans_run: HLLD is 3D");
    if ((engine == CANS_ENGINE_MLW3 || engine == CANS_ENGINE_ROE3) && dim == 3)
        return abi_fail(CANS_ERR_ARG, "cans_run: the three-component engines are 1D/2D");
    if (engine == CANS_ENGINE_MLWC && (dim == 1 || !bx1d || !bxm1d))
        return abi_fail(CANS_ERR_ARG, "cans_run: mlw_m_c is 2D/3D and needs the radii x, xm (in bx1d, bxm1d)");
    if (engine == CANS_ENGINE_MLWS && (dim == 1 || !bx1d))
        return abi_fail(CANS_ERR_ARG, "cans_run: mlw_m_s is 2D/3D and needs x, xm, y, ym (packed in bx1d)");
    if (engine < CANS_ENGINE_MLW || engine > CANS_ENGINE_MLWS) return abi_fail(CANS_ERR_ARG, "cans_run: engine %d", engine);
    if (engine == CANS_ENGINE_MLW && dim == 2 && !bc)
        return cans2d_mlw_run(a, b, clk, step0, nsteps, qav, gm, safety, dx, dxm, ix, dy, dym, jx, nullptr, margin,
                              (double *)ws, events, stream);
    if (!bc) return abi_fail(CANS_ERR_ARG, "cans_run: bc is NULL (only 2D MLW has the fused periodic path)");
    if (dim == 1) { jx = 1; kx = 1; }
    if (dim == 2) kx = 1;
    const bool m3 = engine == CANS_ENGINE_MLW3;
    const bool cyl = engine == CANS_ENGINE_MLWC && dim == 2;   // fields ro pr vx vz bx bz ay
    const bool cyl3 = engine == CANS_ENGINE_MLWC && dim == 3;  // the 3D field list, radii x, xm
    const bool sph = engine == CANS_ENGINE_MLWS;               // x xm y ym packed in bx1d
    const bool c3 = m3 || engine == CANS_ENGINE_ROE3;      // three-component field lists
    const int nf = c3 ? (dim == 1 ? 7 : 9) : (dim == 1 ? 5 : (dim == 2 ? 7 : 8));
    const int lo = (engine == CANS_ENGINE_MLW || m3 || cyl || cyl3 || sph) ? 3 : 5;   // Roe: margin >= 3 (drop-in)
    const int n[3] = {ix, jx, kx};
    for (int d = 0; d < dim; d++)
        if (n[d] < lo || n[d] < 2 * margin + 1) return abi_fail(CANS_ERR_ARG, "cans_run: dimension %d too small", n[d]);
    hipStream_t s = (hipStream_t)stream;
    Clock *c = (Clock *)clk;
    const size_t N = (size_t)ix * jx * kx;
    Carver cv{(double *)ws};
    Metric m[3];
    carve_metric(cv, ix, jx, kx, m);
    const double *gx = sph ? bx1d : nullptr, *gxm = sph ? bx1d + ix : nullptr;
    double *trig = sph ? cv.take(6 * (size_t)jx) : nullptr;
    if (sph && step0 == 0)   // sin, cos, tan of y and ym (mlw_m_s.f:109-112, 243, 274; cfl_m_s.f:26)
        if (int rc = host_trig(bx1d + 2 * ix, bx1d + 2 * ix + jx, jx, trig, s)) return rc;
    if ((engine == CANS_ENGINE_MLW || m3 || cyl || cyl3 || sph) && step0 == 0) {   // the first call fills the metric
        HIPCHK(launch_metric(dx, dxm, ix, m[0], s));
        if (dim >= 2) HIPCHK(launch_metric(dy, dym, jx, m[1], s));
        if (dim == 3) HIPCHK(launch_metric(dz, dzm, kx, m[2], s));
    }
    double *eng = cv.p;
    // In a periodic box (every field periodic on every face) the engine's
    // kernels carry cfl_m and the clock, so a step is the engine and the
    // ghost fill:
    //  * 3D Cartesian MLW (marching kernel), HLLD and FP32-flux Roe (fused z
    //    marches), 2D mlw_m3: the kernels advance the clock and take the next
    //    step's cfl_m; the first step of a call still takes cfl_m of its state
    //    in a pass of its own (the same minimum again when the call before
    //    fused it);
    //  * FP64 3D Roe and 2D Roe (roecfl): the stage-1 x faces take the step's
    //    own cfl_m and the next kernel advances the clock, after the first
    //    step of a call (whose ghost cells need not be images yet), which
    //    keeps the cfl pass and the clock kernel.
    // (allper: every field periodic on every face; the ghost fill is then one
    // pass, launch_bnd_periodic)
    bool allper = margin >= 1 && bc != nullptr;
    for (int k = 0; allper && k < 2 * dim * nf; k++) allper = bc[k] == BD_PER;
    for (int d = 0; allper && d < dim; d++) allper = n[d] - 2 * margin >= margin;
    const bool perio = allper && safety > 0.0;
    const bool per3 = perio && dim == 3 &&
                      ((engine == CANS_ENGINE_MLW && mlw3d_path() != 0) ||
                       (engine == CANS_ENGINE_ROE && (!mixed || roe3d_fused_z(ix, jx, kx))) ||
                       engine == CANS_ENGINE_HLLD);
    const bool roecfl = engine == CANS_ENGINE_ROE && ((per3 && !mixed) || (perio && dim == 2));
    const bool m3cfl = perio && dim == 2 && m3;
    const bool fused = per3 || roecfl || m3cfl;
    for (int st = step0; st < step0 + nsteps; st++) {
        double *const *in = (st & 1) ? b : a;
        double *const *out = (st & 1) ? a : b;
        // map the field list onto the engine slots (ro pr vx vy vz bx by bz)
        double *pin[8] = {nullptr}, *pout[8] = {nullptr};
        if (c3) {   // ro pr vx vy vz (by bz) or (bx by bz), az
            const int map1[7] = {0, 1, 2, 3, 4, 6, 7};
            for (int f = 0; f < (dim == 1 ? 7 : 8); f++) {
                const int slot = dim == 1 ? map1[f] : f;
                pin[slot] = in[f];
                pout[slot] = out[f];
            }
        } else if (dim == 1) {
            const int map[5] = {0, 1, 2, 3, 6};
            for (int f = 0; f < 5; f++) { pin[map[f]] = in[f]; pout[map[f]] = out[f]; }
        } else if (dim == 2) {
            const int map[6] = {0, 1, 2, 3, 5, 6};
            for (int f = 0; f < 6; f++) { pin[map[f]] = in[f]; pout[map[f]] = out[f]; }
        } else {
            for (int f = 0; f < 8; f++) { pin[f] = in[f]; pout[f] = out[f]; }
        }
        // cfl_m of the current state (cfl_m.f for the dimension; the HLLD
        // driver also calls the Gaussian cans3d cfl_m)
        CflArgs ca{};
        ca.ro = pin[0]; ca.pr = pin[1]; ca.vx = pin[2]; ca.vy = pin[3]; ca.vz = pin[4];
        ca.bx = dim == 1 ? bx1d : pin[5]; ca.by = pin[6]; ca.bz = pin[7];   // (3D mlw_m_c: the plain cfl_m)
        ca.dx = dx; ca.dy = dy; ca.dz = dz; ca.safety = safety; ca.gm = gm;
        ca.xr = gx; ca.trig = dim == 3 ? trig : nullptr;   // cfl_m_s for the spherical engine
        ca.ix = ix; ca.jx = jx; ca.kx = kx; ca.dim = dim; ca.clk = c; ca.step = st; ca.comp3 = c3 ? 1 : 0;
        if (!fused || st == step0) HIPCHK(launch_cfl(ca, s));
        if (!fused || (roecfl && st == step0)) HIPCHK(launch_clock_step(c, st, s));
        const int e = st - step0;
        if (events) HIPCHK(hipEventRecord((hipEvent_t)events[2 * e], s));
        if (cyl) {
            Mlw3Args g{};
            for (int f = 0; f < 7; f++) { g.in[f] = in[f]; g.out[f] = out[f]; }
            g.x = bx1d; g.xm = bxm1d; g.dxm = dxm; g.dym = dym; g.mx = m[0]; g.my = m[1];
            g.qav = qav; g.gm = gm; g.clk = c; g.step = st; g.ix = ix; g.jx = jx;
            HIPCHK(launch_mlw_m_c(g, s));
        } else if (m3) {
            Mlw3Args g{};
            for (int f = 0; f < nf; f++) { g.in[f] = in[f]; g.out[f] = out[f]; }
            g.bx = bx1d; g.bxm = bxm1d; g.dxm = dxm; g.dym = dym; g.mx = m[0]; g.my = m[1];
            g.qav = qav; g.gm = gm; g.clk = c; g.step = st; g.ix = ix; g.jx = jx;
            if (m3cfl) { g.dx = dx; g.dy = dy; g.safety = safety; g.per_margin = margin; }
            HIPCHK(launch_mlw_m3(g, dim, s));
        } else if (sph && dim == 2) {
            Mlw2sArgs g{};
            for (int f = 0; f < 7; f++) { g.prim[f] = in[f]; g.out[f] = out[f]; }
            Carver ec{eng};
            g.cell = ec.take((size_t)kMlw2sCellFields * N);
            g.corner = ec.take((size_t)kMlw2sCornerFields * N);
            g.x = gx; g.xm = gxm; g.trig = trig; g.dxm = dxm; g.dym = dym; g.mx = m[0]; g.my = m[1];
            g.qav = qav; g.gm = gm; g.clk = c; g.step = st; g.ix = ix; g.jx = jx;
            HIPCHK(launch_mlw2d_s(g, s));
        } else if (sph) {
            Mlw3dArgs g{};
            for (int q = 0; q < 8; q++) { g.prim[q] = pin[q]; g.out[q] = pout[q]; }
            Carver ec{eng};
            g.cell = ec.take((size_t)kMlw3dCylCellFields * N);
            g.corner = ec.take((size_t)kMlw3dCylCornerFields * N);
            g.x = gx; g.xm = gxm; g.trig = trig;
            g.dxm = dxm; g.dym = dym; g.dzm = dzm; g.mx = m[0]; g.my = m[1]; g.mz = m[2];
            g.qav = qav; g.gm = gm; g.clk = c; g.step = st; g.ix = ix; g.jx = jx; g.kx = kx;
            HIPCHK(launch_mlw3d(g, s));
        } else if (cyl3) {
            Mlw3dArgs g{};
            for (int q = 0; q < 8; q++) { g.prim[q] = pin[q]; g.out[q] = pout[q]; }
            Carver ec{eng};
            g.cell = ec.take((size_t)kMlw3dCylCellFields * N);
            g.corner = ec.take((size_t)kMlw3dCylCornerFields * N);
            g.x = bx1d; g.xm = bxm1d;
            g.dxm = dxm; g.dym = dym; g.dzm = dzm; g.mx = m[0]; g.my = m[1]; g.mz = m[2];
            g.qav = qav; g.gm = gm; g.clk = c; g.step = st; g.ix = ix; g.jx = jx; g.kx = kx;
            HIPCHK(launch_mlw3d(g, s));
        } else if (engine == CANS_ENGINE_MLW) {
            if (dim == 1) {
                Mlw1dArgs g{};
                g.ro = pin[0]; g.pr = pin[1]; g.vx = pin[2]; g.vy = pin[3]; g.by = pin[6];
                g.ro_o = pout[0]; g.pr_o = pout[1]; g.vx_o = pout[2]; g.vy_o = pout[3]; g.by_o = pout[6];
                g.bx = bx1d; g.bxm = bxm1d; g.dx = dx; g.dxm = dxm; g.mx = m[0];
                g.qav = qav; g.gm = gm; g.clk = c; g.step = st; g.ix = ix;
                HIPCHK(launch_mlw1d(g, s));
            } else if (dim == 2) {
                Mlw2dArgs g{};
                g.ro = in[0]; g.pr = in[1]; g.vx = in[2]; g.vy = in[3]; g.bx = in[4]; g.by = in[5]; g.az = in[6];
                g.ro_o = out[0]; g.pr_o = out[1]; g.vx_o = out[2]; g.vy_o = out[3]; g.bx_o = out[4];
                g.by_o = out[5]; g.az_o = out[6];
                g.dx = dx; g.dxm = dxm; g.dy = dy; g.dym = dym; g.mx = m[0]; g.my = m[1];
                g.qav = qav; g.gm = gm; g.clk = c; g.step = st; g.ix = ix; g.jx = jx;
                g.per_margin = 0; g.safety = 0.0;
                g.peek = 1;
                HIPCHK(launch_mlw2d(g, s));
            } else {
                Mlw3dArgs g{};
                for (int q = 0; q < 8; q++) { g.prim[q] = pin[q]; g.out[q] = pout[q]; }
                Carver ec{eng};
                g.cell = ec.take((size_t)kMlw3dCellFields * N);
                g.corner = ec.take((size_t)kMlw3dCornerFields * N);
                g.dxm = dxm; g.dym = dym; g.dzm = dzm; g.mx = m[0]; g.my = m[1]; g.mz = m[2];
                g.qav = qav; g.gm = gm; g.clk = c; g.step = st; g.ix = ix; g.jx = jx; g.kx = kx;
                if (per3) { g.dx = dx; g.dy = dy; g.dz = dz; g.safety = safety; g.per_margin = margin; }
                HIPCHK(launch_mlw3d(g, s));
            }
        } else if (engine == CANS_ENGINE_ROE || engine == CANS_ENGINE_ROE3) {
            RoeArgs r{};
            for (int q = 0; q < 8; q++) { r.prim[q] = pin[q]; r.out[q] = pout[q]; }
            Carver ec{eng};
            roe_scratch(r, ec, N, dim, c3);
            const int faz = c3 ? 8 : 6;            // Az's place in the 2D field list
            r.az = dim == 2 ? in[faz] : nullptr;
            r.out_az = dim == 2 ? out[faz] : nullptr;
            r.bx = bx1d; r.bxm = bxm1d; r.dx = dx; r.dy = dy; r.dz = dz;
            r.gm = gm; r.clk = c; r.step = st; r.ix = ix; r.jx = jx; r.kx = kx;
            if (fused) { r.safety = safety; r.per_margin = margin; r.cfl_now = roecfl && st > step0; }
            HIPCHK(launch_roe(r, dim, mixed != 0, s, c3));
        } else {
            HlldArgs h{};
            for (int q = 0; q < 8; q++) { h.prim[q] = pin[q]; h.out[q] = pout[q]; }
            Carver ec{eng};
            hlld_scratch(h, ec, N);
            h.dx = dx; h.dy = dy; h.dz = dz; h.gm = gm; h.mu = mu;
            h.mfdim[0] = h.mfdim[1] = h.mfdim[2] = 1;
            h.clk = c; h.step = st; h.ix = ix; h.jx = jx; h.kx = kx;
            if (per3) { h.safety = safety; h.per_margin = margin; }
            HIPCHK(launch_hlld(h, mixed != 0, s));
        }
        if (events) HIPCHK(hipEventRecord((hipEvent_t)events[2 * e + 1], s));
        // bnd: x faces, then y, then z; each face: periodic fields both sides
        if (allper) {
            HIPCHK(launch_bnd_periodic(out, nf, ix, jx, kx, dim, margin, s));
            continue;
        }
        for (int dir = 0; dir < dim; dir++) {
            for (int side = 0; side < 2; side++) {
                BndArgs bp{}, bo{};
                for (int f = 0; f < nf; f++) {
                    const int kind = bc[(dir * 2 + side) * nf + f];
                    if (kind == BD_NONE) continue;
                    if (kind == BD_PER) {
                        if (side == 1) continue;
                        bp.q[bp.nf] = out[f]; bp.kind[bp.nf] = kind; bp.nf++;
                    } else {
                        bo.q[bo.nf] = out[f]; bo.kind[bo.nf] = kind; bo.nf++;
                    }
                }
                const double *dm = dir == 0 ? dxm : (dir == 1 ? dym : dzm);
                if (bp.nf) {
                    bp.dir = dir; bp.mbnd = margin; bp.margin = margin; bp.dm = dm;
                    bp.ix = ix; bp.jx = jx; bp.kx = kx;
                    HIPCHK(launch_bnd(bp, s));
                }
                if (bo.nf) {
                    bo.dir = dir; bo.mbnd = side; bo.margin = margin; bo.dm = dm;
                    bo.ix = ix; bo.jx = jx; bo.kx = kx;
                    HIPCHK(launch_bnd(bo, s));
                }
            }
        }
    }
    return 0;
}

}  // extern "C"
