// cans_abi_batch.hip -- extern "C" entry points of the batched 1D tubes
// (declared in include/cans_gpu.h): the run, a native autotuner of the
// persistent kernel's launch shape, and a hipGraph of a whole batch run.
#include <hip/hip_runtime.h>
#include <string.h>

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "cans_common.hpp"
#include "cans_kernels.hpp"
#include "../../include/cans_gpu.h"

namespace cans {
int abi_fail(int code, const char *fmt, ...);
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

int cu_count() {
    static int n = 0;
    if (!n) {
        int dev = 0;
        hipDeviceProp_t p;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess) n = p.multiProcessorCount;
        if (n <= 0) n = 256;
    }
    return n;
}

// Workspace: scratch state (5 fields), 1/dx and 1/dxm, queue counter, and
// for the tuner a copy of the state (5 more fields).
size_t ws_doubles(long long ncells, int tune) { return (5 + 2 + (tune ? 5 : 0)) * al((size_t)ncells) + 64; }

struct BatchIn {
    int ntube;
    const long long *off;
    const int *ix;
    const int *order;
    double *f[5];
    const double *bx, *bxm, *dx, *dxm;
    long long ncells;
    double gm, qav, safety, dtmin, tend;
    int nstop, margin, mixed;
    double *t_out;
    int *ns_out, *merr_out;
    double *ws;
    const int *nstop_tube = nullptr;   // device, per tube (NULL: nstop)
};

// Host check of the per-tube layout (the kernel trusts off/ix): every tube
// lies inside the packed arrays and has at least one interior cell.
int validate(const BatchIn &b) {
    if (b.ntube < 0 || b.ncells < 0 || b.margin < 1 || b.nstop < 0)
        return abi_fail(CANS_ERR_ARG, "cans1d_batch: ntube=%d ncells=%lld margin=%d nstop=%d", b.ntube, b.ncells,
                        b.margin, b.nstop);
    if (b.ntube == 0) return 0;
    for (int q = 0; q < 5; q++)
        if (!b.f[q]) return abi_fail(CANS_ERR_ARG, "cans1d_batch: NULL field %d", q);
    if (!b.off || !b.ix || !b.bx || !b.bxm || !b.dx || !b.dxm || !b.t_out || !b.ns_out || !b.merr_out || !b.ws)
        return abi_fail(CANS_ERR_ARG, "cans1d_batch: NULL argument");
    std::vector<long long> off(b.ntube);
    std::vector<int> ix(b.ntube), order;
    HIPCHK(hipMemcpy(off.data(), b.off, sizeof(long long) * b.ntube, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ix.data(), b.ix, sizeof(int) * b.ntube, hipMemcpyDeviceToHost));
    for (int k = 0; k < b.ntube; k++) {
        if (ix[k] < 2 * b.margin + 1 || ix[k] < 3 || off[k] < 0 || off[k] + ix[k] > b.ncells)
            return abi_fail(CANS_ERR_ARG, "cans1d_batch: tube %d has off=%lld ix=%d (ncells=%lld, margin=%d)", k,
                            off[k], ix[k], b.ncells, b.margin);
    }
    if (b.nstop_tube) {
        std::vector<int> ns(b.ntube);
        HIPCHK(hipMemcpy(ns.data(), b.nstop_tube, sizeof(int) * b.ntube, hipMemcpyDeviceToHost));
        for (int k = 0; k < b.ntube; k++)
            if (ns[k] < 0) return abi_fail(CANS_ERR_ARG, "cans1d_batch: nstop of tube %d is %d", k, ns[k]);
    }
    if (b.order) {
        order.resize(b.ntube);
        HIPCHK(hipMemcpy(order.data(), b.order, sizeof(int) * b.ntube, hipMemcpyDeviceToHost));
        std::vector<char> seen(b.ntube, 0);
        for (int k = 0; k < b.ntube; k++) {
            if (order[k] < 0 || order[k] >= b.ntube || seen[order[k]])
                return abi_fail(CANS_ERR_ARG, "cans1d_batch: order is not a permutation (entry %d = %d)", k, order[k]);
            seen[order[k]] = 1;
        }
    }
    return 0;
}

Batch1dArgs make_args(const BatchIn &b, double *const *state) {
    Batch1dArgs a{};
    a.ntube = b.ntube;
    a.ncells = b.ncells;
    a.off = b.off;
    a.ix = b.ix;
    a.order = b.order;
    const size_t n = al((size_t)b.ncells);
    double *w = b.ws;
    for (int q = 0; q < 5; q++) a.f[q] = state[q];
    for (int q = 0; q < 5; q++) a.g[q] = w + q * n;
    a.di = w + 5 * n;
    a.dim = w + 6 * n;
    a.next = reinterpret_cast<unsigned *>(w + 7 * n);
    a.bx = b.bx; a.bxm = b.bxm; a.dx = b.dx; a.dxm = b.dxm;
    a.gm = b.gm; a.qav = b.qav; a.safety = b.safety; a.dtmin = b.dtmin; a.tend = b.tend; a.t0 = 0.0;
    a.nstop = b.nstop; a.margin = b.margin;
    a.nstop_tube = b.nstop_tube;
    a.t_out = b.t_out; a.ns_out = b.ns_out; a.merr_out = b.merr_out;
    return a;
}

// tuning copy of the state sits after the counter
double *tune_base(const BatchIn &b) { return b.ws + 7 * al((size_t)b.ncells) + 64; }

bool shape_ok(int bt, int bpc) {
    return (bt == 64 || bt == 128 || bt == 256 || bt == 512) && bpc >= 1 && bpc <= 32 && bt * bpc <= 2048;
}

using TuneKey = std::tuple<int, long long, int, int>;
std::map<TuneKey, std::pair<int, int>> &tune_cache() {
    static std::map<TuneKey, std::pair<int, int>> m;
    return m;
}
std::mutex tune_mu;

__global__ void copy5_kernel(double *d0, double *d1, double *d2, double *d3, double *d4, const double *s0,
                             const double *s1, const double *s2, const double *s3, const double *s4, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        d0[i] = s0[i]; d1[i] = s1[i]; d2[i] = s2[i]; d3[i] = s3[i]; d4[i] = s4[i];
    }
}

hipError_t copy5(double *const *dst, const double *const *src, long long n, hipStream_t s) {
    hipLaunchKernelGGL(copy5_kernel, dim3(2048), dim3(256), 0, s, dst[0], dst[1], dst[2], dst[3], dst[4], src[0],
                       src[1], src[2], src[3], src[4], n);
    return hipGetLastError();
}

}  // namespace

extern "C" {

size_t cans1d_batch_ws_bytes(long long ncells, int tune) { return ws_doubles(ncells, tune) * sizeof(double); }

int cans1d_batch_run(int ntube, const long long *off, const int *ix, const int *order, double *ro, double *pr,
                     double *vx, double *vy, double *by, const double *bx, const double *bxm, const double *dx,
                     const double *dxm, long long ncells, double gm, double qav, double safety, double dtmin,
                     double tend, int nstop, int margin, int mixed, double *t_out, int *ns_out, int *merr_out,
                     int block_threads, int blocks_per_cu, void *ws, void *stream) {
    BatchIn b{ntube, off, ix, order, {ro, pr, vx, vy, by}, bx, bxm, dx, dxm, ncells, gm, qav, safety, dtmin,
              tend, nstop, margin, mixed, t_out, ns_out, merr_out, (double *)ws};
    if (int e = validate(b)) return e;
    if (ntube == 0) return 0;
    if (!shape_ok(block_threads, blocks_per_cu))
        return abi_fail(CANS_ERR_ARG, "cans1d_batch_run: block_threads=%d blocks_per_cu=%d", block_threads,
                        blocks_per_cu);
    const Batch1dArgs a = make_args(b, b.f);
    HIPCHK(launch_batch1d(a, mixed, block_threads, cu_count() * blocks_per_cu, (hipStream_t)stream));
    return 0;
}

int cans1d_batch_run_steps(int ntube, const long long *off, const int *ix, const int *order, double *ro, double *pr,
                           double *vx, double *vy, double *by, const double *bx, const double *bxm, const double *dx,
                           const double *dxm, long long ncells, double gm, double qav, double safety, double dtmin,
                           double tend, const int *nstop_tube, int margin, int mixed, double *t_out, int *ns_out,
                           int *merr_out, int block_threads, int blocks_per_cu, void *ws, void *stream) {
    if (!nstop_tube) return abi_fail(CANS_ERR_ARG, "cans1d_batch_run_steps: NULL nstop_tube");
    BatchIn b{ntube, off, ix, order, {ro, pr, vx, vy, by}, bx, bxm, dx, dxm, ncells, gm, qav, safety, dtmin,
              tend, 0, margin, mixed, t_out, ns_out, merr_out, (double *)ws, nstop_tube};
    if (int e = validate(b)) return e;
    if (ntube == 0) return 0;
    if (!shape_ok(block_threads, blocks_per_cu))
        return abi_fail(CANS_ERR_ARG, "cans1d_batch_run_steps: block_threads=%d blocks_per_cu=%d", block_threads,
                        blocks_per_cu);
    const Batch1dArgs a = make_args(b, b.f);
    HIPCHK(launch_batch1d(a, mixed, block_threads, cu_count() * blocks_per_cu, (hipStream_t)stream));
    return 0;
}

int cans1d_batch_tune(int ntube, const long long *off, const int *ix, const int *order, double *ro, double *pr,
                      double *vx, double *vy, double *by, const double *bx, const double *bxm, const double *dx,
                      const double *dxm, long long ncells, double gm, double qav, double safety, double dtmin,
                      int margin, int mixed, int probe_steps, double *t_out, int *ns_out, int *merr_out, void *ws,
                      int *block_threads, int *blocks_per_cu, float *best_ms, void *stream) {
    BatchIn b{ntube, off, ix, order, {ro, pr, vx, vy, by}, bx, bxm, dx, dxm, ncells, gm, qav, safety, dtmin,
              1e300, probe_steps, margin, mixed, t_out, ns_out, merr_out, (double *)ws};
    if (!block_threads || !blocks_per_cu) return abi_fail(CANS_ERR_ARG, "cans1d_batch_tune: NULL output");
    if (probe_steps < 1) return abi_fail(CANS_ERR_ARG, "cans1d_batch_tune: probe_steps=%d", probe_steps);
    if (int e = validate(b)) return e;
    const TuneKey key{ntube, ncells, mixed, probe_steps};
    {
        std::lock_guard<std::mutex> g(tune_mu);
        auto it = tune_cache().find(key);
        if (it != tune_cache().end()) {
            *block_threads = it->second.first;
            *blocks_per_cu = it->second.second;
            if (best_ms) *best_ms = -1.0f;
            return 0;
        }
    }
    hipStream_t s = (hipStream_t)stream;
    // The probes run on a copy of the state, probe_steps steps per tube.
    double *tc = tune_base(b);
    const size_t n = al((size_t)ncells);
    double *copy[5] = {tc, tc + n, tc + 2 * n, tc + 3 * n, tc + 4 * n};
    const double *src[5] = {ro, pr, vx, vy, by};
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const int bts[] = {64, 128, 256, 512};
    const int bpcs[] = {1, 2, 4, 8, 16};
    float best = 1e30f;
    int bbt = 256, bbpc = 2;
    for (int bt : bts) {
        for (int bpc : bpcs) {
            if (!shape_ok(bt, bpc)) continue;
            float ms = 0.0f;
            for (int rep = 0; rep < 2; rep++) {   // the first run warms caches and code
                HIPCHK(copy5(copy, src, ncells, s));
                const Batch1dArgs a = make_args(b, copy);
                HIPCHK(hipEventRecord(e0, s));
                HIPCHK(launch_batch1d(a, mixed, bt, cu_count() * bpc, s));
                HIPCHK(hipEventRecord(e1, s));
                HIPCHK(hipEventSynchronize(e1));
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            }
            if (ms < best) { best = ms; bbt = bt; bbpc = bpc; }
        }
    }
    HIPCHK(hipEventDestroy(e0));
    HIPCHK(hipEventDestroy(e1));
    {
        std::lock_guard<std::mutex> g(tune_mu);
        tune_cache()[key] = {bbt, bbpc};
    }
    *block_threads = bbt;
    *blocks_per_cu = bbpc;
    if (best_ms) *best_ms = best;
    return 0;
}

int cans1d_batch_graph(int ntube, const long long *off, const int *ix, const int *order, double *ro, double *pr,
                       double *vx, double *vy, double *by, const double *bx, const double *bxm, const double *dx,
                       const double *dxm, long long ncells, double gm, double qav, double safety, double dtmin,
                       double tend, int nstop, int margin, int mixed, double *t_out, int *ns_out, int *merr_out,
                       int block_threads, int blocks_per_cu, void *ws, const double *const *init5, void **graph) {
    BatchIn b{ntube, off, ix, order, {ro, pr, vx, vy, by}, bx, bxm, dx, dxm, ncells, gm, qav, safety, dtmin,
              tend, nstop, margin, mixed, t_out, ns_out, merr_out, (double *)ws};
    if (!graph) return abi_fail(CANS_ERR_ARG, "cans1d_batch_graph: NULL graph");
    if (int e = validate(b)) return e;
    if (ntube == 0) return abi_fail(CANS_ERR_ARG, "cans1d_batch_graph: empty batch");
    if (!shape_ok(block_threads, blocks_per_cu))
        return abi_fail(CANS_ERR_ARG, "cans1d_batch_graph: block_threads=%d blocks_per_cu=%d", block_threads,
                        blocks_per_cu);
    hipStream_t s;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    if (init5) e = copy5(b.f, init5, ncells, s);   // each replay starts from init5
    if (e == hipSuccess) e = launch_batch1d(make_args(b, b.f), mixed, block_threads, cu_count() * blocks_per_cu, s);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(s, &g);
    if (e == hipSuccess) e = e2;
    hipGraphExec_t x = nullptr;
    if (e == hipSuccess) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    (void)hipStreamDestroy(s);
    if (e != hipSuccess) return abi_fail(CANS_ERR_HIP, "cans1d_batch_graph: %s", hipGetErrorString(e));
    *graph = x;
    return 0;
}

int cans_graph_launch(void *graph, void *stream) {
    if (!graph) return abi_fail(CANS_ERR_ARG, "cans_graph_launch: NULL graph");
    HIPCHK(hipGraphLaunch((hipGraphExec_t)graph, (hipStream_t)stream));
    return 0;
}

int cans_graph_destroy(void *graph) {
    if (graph) HIPCHK(hipGraphExecDestroy((hipGraphExec_t)graph));
    return 0;
}

}  // extern "C"

// ======================================================================
// Batched periodic 2D runs (MLW)
namespace {

// workspace: the scratch state (7 fields), the x and y metric (4 arrays
// each), the per-run argument table (2 per run), the queue counter, and
// (tune) a copy of the state for the tuner's probes
size_t b2_table_bytes(int nrun) { return (2 * (size_t)nrun * sizeof(Mlw2dArgs) + 255) & ~size_t(255); }
size_t b2_ws_bytes(long long ncells, long long nx1d, long long ny1d, int nrun, int tune) {
    return (7 * al((size_t)ncells) + 4 * al((size_t)nx1d) + 4 * al((size_t)ny1d)) * sizeof(double) +
           b2_table_bytes(nrun) + 256 + (tune ? 7 * al((size_t)ncells) * sizeof(double) : 0);
}
double *b2_tune_copy(void *ws, long long ncells, long long nx1d, long long ny1d, int nrun) {
    return (double *)((char *)ws + b2_ws_bytes(ncells, nx1d, ny1d, nrun, 0));
}

__global__ void copy7_kernel(Batch2dArgs dst, Batch2dArgs src, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
#pragma unroll
        for (int q = 0; q < 7; q++) dst.f[q][i] = src.f[q][i];
}
hipError_t copy7(double *const *dst, const double *const *src, long long n, hipStream_t s) {
    Batch2dArgs d{}, c{};
    for (int q = 0; q < 7; q++) {
        d.f[q] = dst[q];
        c.f[q] = const_cast<double *>(src[q]);
    }
    hipLaunchKernelGGL(copy7_kernel, dim3(2048), dim3(256), 0, s, d, c, n);
    return hipGetLastError();
}

bool b2_shape_ok(int bt, int bpc) { return shape_ok(bt, bpc) && bt * bpc <= 512; }

using Tune2Key = std::tuple<int, long long, long long, int>;
std::map<Tune2Key, std::pair<int, int>> &tune2_cache() {
    static std::map<Tune2Key, std::pair<int, int>> m;
    return m;
}

// Checks the layout on the host (the kernel trusts it), fills the launch
// arguments for the state `state7` and uploads the per-run argument table
// (synchronously: the host copy must outlive the transfer).
int b2_prepare(const char *who, int nrun, const long long *off, const int *ix, const int *jx, const long long *offx,
               const long long *offy, const int *order, double *const *state7, const double *dx, const double *dxm,
               const double *dy, const double *dym, long long ncells, long long nx1d, long long ny1d, double gm,
               double qav, double safety, double dtmin, double tend, int nstop, int margin, double *t_out,
               int *ns_out, int *merr_out, void *ws, hipStream_t s, Batch2dArgs &b) {
    if (nrun < 0 || ncells < 0 || nx1d < 0 || ny1d < 0 || margin < 1 || nstop < 0)
        return abi_fail(CANS_ERR_ARG, "%s: nrun=%d ncells=%lld margin=%d nstop=%d", who, nrun, ncells, margin, nstop);
    if (!state7 || !off || !ix || !jx || !offx || !offy || !dx || !dxm || !dy || !dym || !t_out || !ns_out ||
        !merr_out || !ws)
        return abi_fail(CANS_ERR_ARG, "%s: NULL argument", who);
    for (int q = 0; q < 7; q++)
        if (!state7[q]) return abi_fail(CANS_ERR_ARG, "%s: NULL field %d", who, q);
    std::vector<long long> o(nrun), ox(nrun), oy(nrun);
    std::vector<int> nx(nrun), ny(nrun);
    if (nrun > 0) {
        HIPCHK(hipMemcpy(o.data(), off, sizeof(long long) * nrun, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(ox.data(), offx, sizeof(long long) * nrun, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(oy.data(), offy, sizeof(long long) * nrun, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(nx.data(), ix, sizeof(int) * nrun, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(ny.data(), jx, sizeof(int) * nrun, hipMemcpyDeviceToHost));
    }
    for (int k = 0; k < nrun; k++) {
        const bool ok = nx[k] >= 3 && ny[k] >= 3 && nx[k] - 2 * margin >= margin && ny[k] - 2 * margin >= margin &&
                        o[k] >= 0 && o[k] + (long long)nx[k] * ny[k] <= ncells && ox[k] >= 0 &&
                        ox[k] + nx[k] <= nx1d && oy[k] >= 0 && oy[k] + ny[k] <= ny1d;
        if (!ok)
            return abi_fail(CANS_ERR_ARG, "%s: run %d has ix=%d jx=%d off=%lld offx=%lld offy=%lld "
                            "(ncells=%lld nx1d=%lld ny1d=%lld margin=%d; the periodic box needs n-2*margin >= margin)",
                            who, k, nx[k], ny[k], o[k], ox[k], oy[k], ncells, nx1d, ny1d, margin);
    }
    if (order && nrun > 0) {
        std::vector<int> ord(nrun);
        HIPCHK(hipMemcpy(ord.data(), order, sizeof(int) * nrun, hipMemcpyDeviceToHost));
        std::vector<char> seen(nrun, 0);
        for (int k = 0; k < nrun; k++) {
            if (ord[k] < 0 || ord[k] >= nrun || seen[ord[k]])
                return abi_fail(CANS_ERR_ARG, "%s: order is not a permutation (entry %d = %d)", who, k, ord[k]);
            seen[ord[k]] = 1;
        }
    }
    b = Batch2dArgs{};
    b.nrun = nrun;
    b.off = off; b.offx = offx; b.offy = offy; b.ix = ix; b.jx = jx; b.order = order;
    double *w = (double *)ws;
    const size_t nc = al((size_t)ncells), n1x = al((size_t)nx1d), n1y = al((size_t)ny1d);
    for (int q = 0; q < 7; q++) {
        b.f[q] = state7[q];
        b.g[q] = w + q * nc;
    }
    for (int q = 0; q < 4; q++) {
        b.mx[q] = w + 7 * nc + q * n1x;
        b.my[q] = w + 7 * nc + 4 * n1x + q * n1y;
    }
    Mlw2dArgs *table = reinterpret_cast<Mlw2dArgs *>(w + 7 * nc + 4 * n1x + 4 * n1y);
    b.next = reinterpret_cast<unsigned *>(reinterpret_cast<char *>(table) + b2_table_bytes(nrun));
    b.runs = table;
    b.dx = dx; b.dxm = dxm; b.dy = dy; b.dym = dym;
    b.gm = gm; b.qav = qav; b.safety = safety; b.dtmin = dtmin; b.tend = tend;
    b.nstop = nstop; b.margin = margin;
    b.t_out = t_out; b.ns_out = ns_out; b.merr_out = merr_out;
    if (nrun == 0) return 0;
    // the per-run arguments of the strip march, one per step parity
    std::vector<Mlw2dArgs> h(2 * (size_t)nrun);
    for (int k = 0; k < nrun; k++)
        for (int par = 0; par < 2; par++) {
            Mlw2dArgs &a = h[2 * (size_t)k + par];
            a = Mlw2dArgs{};
            double *const *cur = par ? b.g : b.f;
            double *const *nxt = par ? b.f : b.g;
            a.ro = cur[0] + o[k]; a.pr = cur[1] + o[k]; a.vx = cur[2] + o[k]; a.vy = cur[3] + o[k];
            a.bx = cur[4] + o[k]; a.by = cur[5] + o[k]; a.az = cur[6] + o[k];
            a.ro_o = nxt[0] + o[k]; a.pr_o = nxt[1] + o[k]; a.vx_o = nxt[2] + o[k]; a.vy_o = nxt[3] + o[k];
            a.bx_o = nxt[4] + o[k]; a.by_o = nxt[5] + o[k]; a.az_o = nxt[6] + o[k];
            a.dx = dx + ox[k]; a.dxm = dxm + ox[k]; a.dy = dy + oy[k]; a.dym = dym + oy[k];
            a.mx = Metric{b.mx[0] + ox[k], b.mx[1] + ox[k], b.mx[2] + ox[k], b.mx[3] + ox[k]};
            a.my = Metric{b.my[0] + oy[k], b.my[1] + oy[k], b.my[2] + oy[k], b.my[3] + oy[k]};
            a.qav = qav; a.gm = gm; a.dt = 0.0; a.clk = nullptr; a.step = 0;
            a.ix = nx[k]; a.jx = ny[k];
            a.per_margin = margin; a.safety = safety; a.peek = 0; a.dt_log = nullptr;
        }
    // ordered after any earlier work on the stream that reads the table
    HIPCHK(hipMemcpyAsync(table, h.data(), h.size() * sizeof(Mlw2dArgs), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

}  // namespace

extern "C" {

size_t cans2d_batch_ws_bytes(long long ncells, long long nx1d, long long ny1d, int nrun, int tune) {
    return b2_ws_bytes(ncells, nx1d, ny1d, nrun, tune);
}

int cans2d_batch_run(int nrun, const long long *off, const int *ix, const int *jx, const long long *offx,
                     const long long *offy, const int *order, double *const *state7, const double *dx,
                     const double *dxm, const double *dy, const double *dym, long long ncells, long long nx1d,
                     long long ny1d, double gm, double qav, double safety, double dtmin, double tend, int nstop,
                     int margin, double *t_out, int *ns_out, int *merr_out, int block_threads, int blocks_per_cu,
                     void *ws, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    if (!b2_shape_ok(block_threads, blocks_per_cu))
        return abi_fail(CANS_ERR_ARG, "cans2d_batch_run: block_threads=%d blocks_per_cu=%d (at most 512 threads "
                        "per CU: two waves per SIMD)", block_threads, blocks_per_cu);
    Batch2dArgs b;
    if (int e = b2_prepare("cans2d_batch_run", nrun, off, ix, jx, offx, offy, order, state7, dx, dxm, dy, dym, ncells,
                           nx1d, ny1d, gm, qav, safety, dtmin, tend, nstop, margin, t_out, ns_out, merr_out, ws, s, b))
        return e;
    if (nrun == 0) return 0;
    HIPCHK(launch_batch2d(b, block_threads, cu_count() * blocks_per_cu, s));
    return 0;
}

int cans2d_batch_tune(int nrun, const long long *off, const int *ix, const int *jx, const long long *offx,
                      const long long *offy, const int *order, double *const *state7, const double *dx,
                      const double *dxm, const double *dy, const double *dym, long long ncells, long long nx1d,
                      long long ny1d, double gm, double qav, double safety, double dtmin, int margin, int probe_steps,
                      double *t_out, int *ns_out, int *merr_out, void *ws, int *block_threads, int *blocks_per_cu,
                      float *best_ms, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    if (!block_threads || !blocks_per_cu) return abi_fail(CANS_ERR_ARG, "cans2d_batch_tune: NULL output");
    if (probe_steps < 1) return abi_fail(CANS_ERR_ARG, "cans2d_batch_tune: probe_steps=%d", probe_steps);
    if (!state7 || !ws) return abi_fail(CANS_ERR_ARG, "cans2d_batch_tune: NULL argument");
    // the probes run on a copy of the state (room from cans2d_batch_ws_bytes(..., tune=1))
    double *tc = b2_tune_copy(ws, ncells, nx1d, ny1d, nrun);
    const size_t nc = al((size_t)ncells);
    double *copy[7];
    for (int q = 0; q < 7; q++) copy[q] = tc + q * nc;
    Batch2dArgs b;
    if (int e = b2_prepare("cans2d_batch_tune", nrun, off, ix, jx, offx, offy, order, copy, dx, dxm, dy, dym, ncells,
                           nx1d, ny1d, gm, qav, safety, dtmin, 1e300, probe_steps, margin, t_out, ns_out, merr_out,
                           ws, s, b))
        return e;
    const Tune2Key key{nrun, ncells, nx1d + ny1d, probe_steps};
    {
        std::lock_guard<std::mutex> g(tune_mu);
        auto it = tune2_cache().find(key);
        if (it != tune2_cache().end()) {
            *block_threads = it->second.first;
            *blocks_per_cu = it->second.second;
            if (best_ms) *best_ms = -1.0f;
            return 0;
        }
    }
    if (nrun == 0) {
        *block_threads = 512;
        *blocks_per_cu = 1;
        if (best_ms) *best_ms = 0.0f;
        return 0;
    }
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const int bts[] = {64, 128, 256, 512};
    const int bpcs[] = {1, 2, 4, 8};
    float best = 1e30f;
    int bbt = 512, bbpc = 1;
    for (int bt : bts) {
        for (int bpc : bpcs) {
            if (!b2_shape_ok(bt, bpc)) continue;
            float ms = 0.0f;
            for (int rep = 0; rep < 2; rep++) {   // the first run warms caches and code
                HIPCHK(copy7(copy, state7, ncells, s));
                HIPCHK(hipEventRecord(e0, s));
                HIPCHK(launch_batch2d(b, bt, cu_count() * bpc, s));
                HIPCHK(hipEventRecord(e1, s));
                HIPCHK(hipEventSynchronize(e1));
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            }
            if (ms < best) { best = ms; bbt = bt; bbpc = bpc; }
        }
    }
    HIPCHK(hipEventDestroy(e0));
    HIPCHK(hipEventDestroy(e1));
    {
        std::lock_guard<std::mutex> g(tune_mu);
        tune2_cache()[key] = {bbt, bbpc};
    }
    *block_threads = bbt;
    *blocks_per_cu = bbpc;
    if (best_ms) *best_ms = best;
    return 0;
}

int cans2d_batch_graph(int nrun, const long long *off, const int *ix, const int *jx, const long long *offx,
                       const long long *offy, const int *order, double *const *state7, const double *dx,
                       const double *dxm, const double *dy, const double *dym, long long ncells, long long nx1d,
                       long long ny1d, double gm, double qav, double safety, double dtmin, double tend, int nstop,
                       int margin, double *t_out, int *ns_out, int *merr_out, int block_threads, int blocks_per_cu,
                       void *ws, const double *const *init7, void **graph) {
    if (!graph) return abi_fail(CANS_ERR_ARG, "cans2d_batch_graph: NULL graph");
    if (nrun == 0) return abi_fail(CANS_ERR_ARG, "cans2d_batch_graph: empty batch");
    if (!b2_shape_ok(block_threads, blocks_per_cu))
        return abi_fail(CANS_ERR_ARG, "cans2d_batch_graph: block_threads=%d blocks_per_cu=%d", block_threads,
                        blocks_per_cu);
    if (init7)
        for (int q = 0; q < 7; q++)
            if (!init7[q]) return abi_fail(CANS_ERR_ARG, "cans2d_batch_graph: NULL init field %d", q);
    hipStream_t s;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    Batch2dArgs b;
    // the table is uploaded before the capture; replays read it from the workspace
    if (int e = b2_prepare("cans2d_batch_graph", nrun, off, ix, jx, offx, offy, order, state7, dx, dxm, dy, dym,
                           ncells, nx1d, ny1d, gm, qav, safety, dtmin, tend, nstop, margin, t_out, ns_out, merr_out,
                           ws, s, b)) {
        (void)hipStreamDestroy(s);
        return e;
    }
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        if (init7) e = copy7(state7, init7, ncells, s);   // each replay starts from init7
        if (e == hipSuccess) e = launch_batch2d(b, block_threads, cu_count() * blocks_per_cu, s);
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(s, &g);
        if (e == hipSuccess) e = e2;
        hipGraphExec_t x = nullptr;
        if (e == hipSuccess) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (e == hipSuccess) *graph = x;
    }
    (void)hipStreamDestroy(s);
    if (e != hipSuccess) return abi_fail(CANS_ERR_HIP, "cans2d_batch_graph: %s", hipGetErrorString(e));
    return 0;
}

}  // extern "C"
