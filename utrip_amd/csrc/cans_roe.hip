// cans_roe.hip -- Roe + MUSCL-minmod MHD engines (cans{1,2,3}d/hdroe/roe_m.f)
// on gfx950, FP64 (bit-exact) and FP32-flux mixed precision.
//
// One step = first-order Roe fluxes on every face of each direction, the
// half-step predictor U^{n+1/2} = U + 1/2 dt sum_d (F_{i-1}-F_i)/dx, MUSCL-
// minmod reconstruction of the half-step primitives, Roe fluxes again, and
// the corrector U^{n+1} = U + dt sum_d (F_{i-1}-F_i)/dx.  Roe is FP64-VALU
// bound (~500-700 flops per face), so each face flux is computed exactly
// once (a flux kernel per direction writing a face array), and the per-cell
// kernels only combine fluxes.  Arrays are SoA FP64, column-major (the face
// arrays of an FP32-flux run hold FP32).
//
// Fields: U  = ro, ee, rx, ry, rz, bx, by, bz (8; 2D: no rz/bz; 1D: ro ee rx ry by)
//         P  = ro, pr, vx, vy, vz, bx, by, bz (primitives, same slots)
// C3 (1D/2D): roe_m3 (cans{1,2}d/hdroe/roe_m3.f), the same steps with the
// z components and the roeflux_m3 flux; its 2D y sweep uses the frame
// (y, x, z) (roe_m3.f:124-125), not the cyclic (y, z, x) of 3D.
// In a periodic run the step carries cfl_m and the clock in its own
// kernels (RoeArgs::per_margin): the FP32-flux 3D step takes the next step's
// term in its stage-2 z march, the FP64 3D and the 2D steps take the
// step's own term in their stage-1 x faces.
#include "cans_common.hpp"
#include "cans_clock.hpp"
#include "cans_kernels.hpp"
#include "cans_roe.hpp"
#include "cans_roeflux.hpp"

namespace cans {
namespace {

// slot layout shared by U and P
enum { S_RO = 0, S_E = 1, S_X = 2, S_Y = 3, S_Z = 4, S_BX = 5, S_BY = 6, S_BZ = 7 };

CANS_DEV size_t lin(int i, int j, int k, int ix, int jx) {
    return (size_t)(i - 1) + (size_t)ix * ((size_t)(j - 1) + (size_t)jx * (size_t)(k - 1));
}

template <int D> struct RoePi2 {};


// Run flag / dt of the step: from the clock (after the step's clock kernel
// ran, or, with `enter`, advanced here by the step's first kernel, block 0
// writing it), or the by-value dt of a drop-in call.
CANS_DEV bool step_dt(const RoeArgs &a, double &dt, bool enter = false) {
    if (!a.clk) { dt = a.dtp ? *a.dtp : a.dt; return true; }
    const StepCtl c = clock_use(a.clk, a.step, !enter, blockIdx.x == 0 && threadIdx.x == 0);
    dt = c.dt;
    return c.run;
}

// Face fluxes are stored in the type they are computed in: FP32-flux runs
// keep FP32 face arrays (half the scratch traffic); widening a float to
// double is exact, so the cell kernels see the same values either way.
template <class T> CANS_DEV void fput(double *f, size_t p, T v) { reinterpret_cast<T *>(f)[p] = v; }
template <class T> CANS_DEV double fget(const double *f, size_t p) { return double(reinterpret_cast<const T *>(f)[p]); }

// ------------------------------------------------------------------ fluxes
// The face of thread t of roe_flux_kernel; with CFL, cq = the cfl_m term
// of the face's left cell when that cell is interior (else left as it is).
template <int D, class T, bool C3, bool CFL>
CANS_DEV void roe_face(const RoeArgs &a, int dir, int stage, const int *lo, const int *cnt, size_t t, double &cq) {
    constexpr bool FULL = D == 3 || C3;      // all three components present
    const int ix = a.ix, jx = a.jx, kx = a.kx;
    const int i = lo[0] + (int)(t % cnt[0]);
    const int j = lo[1] + (int)((t / cnt[0]) % cnt[1]);
    const int k = lo[2] + (int)(t / ((size_t)cnt[0] * cnt[1]));
    const size_t p = lin(i, j, k, ix, jx);
    const size_t s = dir == 0 ? 1 : (dir == 1 ? (size_t)ix : (size_t)ix * jx);
    const double *const *P = stage == 1 ? a.prim : a.half;
    // left / right states of the needed slots
    constexpr int NS = FULL ? (D == 1 ? 7 : 8) : (D == 2 ? 6 : 5);
    // slots used: 1D: ro pr vx vy by ; 2D: + bx ; 3D: all ; 1D C3: all but bx
    const int slots3[8] = {S_RO, S_E, S_X, S_Y, S_Z, S_BX, S_BY, S_BZ};
    const int slots2[6] = {S_RO, S_E, S_X, S_Y, S_BX, S_BY};
    const int slots1[5] = {S_RO, S_E, S_X, S_Y, S_BY};
    const int slots13[7] = {S_RO, S_E, S_X, S_Y, S_Z, S_BY, S_BZ};
    T L[8], R[8];
#pragma unroll
    for (int q = 0; q < 8; q++) { L[q] = T(0); R[q] = T(0); }
#pragma unroll
    for (int m = 0; m < NS; m++) {
        const int q = FULL ? (D == 1 ? slots13[m] : slots3[m]) : (D == 2 ? slots2[m] : slots1[m]);
        const double *da = P[q];
        if (stage == 1) {
            L[q] = T(da[p]);
            R[q] = T(da[p + s]);
        } else if (D == 3) {   // cans3d/hdroe/tvdminmod.f:15-55
            const double d0 = da[p - s], d1 = da[p], d2 = da[p + s], d3 = da[p + 2 * s];
            L[q] = T(d1 + 0.5 * flmt(d1 - d0, d2 - d1));
            R[q] = T(d2 - 0.5 * flmt(d3 - d2, d2 - d1));
        } else {               // cans{1,2}d/hdroe/tvdminmod.f
            const double d0 = da[p - s], d1 = da[p], d2 = da[p + s], d3 = da[p + 2 * s];
            L[q] = T(d1 + 0.5 * flmt(d2 - d1, d1 - d0));
            R[q] = T(d2 - 0.5 * flmt(d2 - d1, d3 - d2));
        }
    }
    if (D == 1) {   // the face field is bxm(i) on both sides in both stages
        L[S_BX] = T(a.bxm[i - 1]);
        R[S_BX] = T(a.bxm[i - 1]);
    }
    if (CFL && D == 3) {   // cans3d cfl_m.f:45-54 of cell i (the stage-1 left state is the cell's own, FP64)
        const int mg = a.per_margin;
        if (i >= mg + 1 && i <= ix - mg && j >= mg + 1 && j <= jx - mg && k >= mg + 1 && k <= kx - mg) {
            const double v2 = L[S_X] * L[S_X] + L[S_Y] * L[S_Y] + L[S_Z] * L[S_Z];
            const double b2 = L[S_BX] * L[S_BX] + L[S_BY] * L[S_BY] + L[S_BZ] * L[S_BZ];
            const double dg = fminF(fminF(image_min(a.dx, i, ix, mg), image_min(a.dy, j, jx, mg)),
                                    image_min(a.dz, k, kx, mg));
            cq = cfl3_term(a.safety, a.gm, L[S_RO], L[S_E], v2, b2, dg);
        }
    }
    if (CFL && D == 2) {   // cans2d cfl_m.f:46-55 of cell i from its FP64 primitives
        const int mg = a.per_margin;
        if (i >= mg + 1 && i <= ix - mg && j >= mg + 1 && j <= jx - mg) {
            const double *const *Q = a.prim;
            const double vx = Q[S_X][p], vy = Q[S_Y][p], bx = Q[S_BX][p], by = Q[S_BY][p];
            const double v2 = vx * vx + vy * vy;
            const double b2 = bx * bx + by * by;
            const double dg = fminF(image_min(a.dx, i, ix, mg), image_min(a.dy, j, jx, mg));
            cq = cfl2_term(a.safety, a.gm, Q[S_RO][p], Q[S_E][p], v2, b2, dg);
        }
    }
    const T gm = T(a.gm);
    double *const *F = a.flux[dir];
    if (D == 3) {
        // sweep frame: (n, t1, t2) = (x,y,z), (y,z,x), (z,x,y)
        const int vn = S_X + dir, v1 = S_X + (dir + 1) % 3, v2 = S_X + (dir + 2) % 3;
        const int bn = S_BX + dir, b1 = S_BX + (dir + 1) % 3, b2 = S_BX + (dir + 2) % 3;
        T o[7];
        roeflux3<T>(gm, L[S_RO], L[S_E], L[vn], L[v1], L[v2], L[bn], L[b1], L[b2], R[S_RO], R[S_E], R[vn], R[v1],
                    R[v2], R[bn], R[b1], R[b2], o);
        fput<T>(F[S_RO], p, o[0]); fput<T>(F[S_E], p, o[1]);
        fput<T>(F[vn], p, o[2]); fput<T>(F[v1], p, o[3]); fput<T>(F[v2], p, o[4]);
        fput<T>(F[b1], p, o[5]); fput<T>(F[b2], p, o[6]);
    } else if (C3) {
        // roe_m3.f: x = (x, y, z); 2D y = (y, x, z)
        const int vn = dir == 0 ? S_X : S_Y, v1 = dir == 0 ? S_Y : S_X;
        const int bn = dir == 0 ? S_BX : S_BY, b1 = dir == 0 ? S_BY : S_BX;
        T o[7];
        roeflux3<T, false>(gm, L[S_RO], L[S_E], L[vn], L[v1], L[S_Z], L[bn], L[b1], L[S_BZ], R[S_RO], R[S_E], R[vn],
                           R[v1], R[S_Z], R[bn], R[b1], R[S_BZ], o);
        fput<T>(F[S_RO], p, o[0]); fput<T>(F[S_E], p, o[1]);
        fput<T>(F[vn], p, o[2]); fput<T>(F[v1], p, o[3]); fput<T>(F[S_Z], p, o[4]);
        fput<T>(F[b1], p, o[5]); fput<T>(F[S_BZ], p, o[6]);
    } else {
        // 2D x: (vx,vy,bx,by) -> ro ee rx ry by ; 2D y: (vy,vx,by,bx) -> ro ee ry rx bx ; 1D as 2D x
        const int vn = dir == 0 ? S_X : S_Y, vt = dir == 0 ? S_Y : S_X;
        const int bn = dir == 0 ? S_BX : S_BY, bt = dir == 0 ? S_BY : S_BX;
        T o[5];
        roeflux12<T>(gm, L[S_RO], L[S_E], L[vn], L[vt], L[bn], L[bt], R[S_RO], R[S_E], R[vn], R[vt], R[bn], R[bt], o);
        fput<T>(F[S_RO], p, o[0]); fput<T>(F[S_E], p, o[1]);
        fput<T>(F[vn], p, o[2]); fput<T>(F[vt], p, o[3]); fput<T>(F[bt], p, o[4]);
    }
}

// Face (i,j,k) in direction `dir` lies between cell (i,j,k) and the next
// cell along `dir`.  stage 1: first-order states from P; stage 2: MUSCL-
// minmod states from the half-step primitives Ph.  The face range is what
// the following cell kernel reads: along dir [lo_n, n-1 or n-2], across
// [lo_t, n+1-lo_t].
// CFL (3D FP64, stage-1 x faces of a periodic run, not the first step of a
// call): the kernel also takes this step's cfl_m over the interior cells it
// reads as the left cell of a face (the state's ghost cells are images from
// the previous step's bnd; image_min, cfl3_term) -- the stage-1 fluxes do
// not need dt -- and the stage-1 y march that follows advances the clock.
template <int D, class T, bool C3, bool CFL = false>
__global__ __launch_bounds__(256) void roe_flux_kernel(RoeArgs a, int dir, int stage) {
    __shared__ double red[4];
    double dtv;
    if (CFL) {
        if (!clock_live(a.clk, a.step)) return;   // uniform across the launch
    } else if (!step_dt(a, dtv, (D == 3 && sizeof(T) == 4 && a.per_margin > 0 && stage == 1 && dir == 0) ||
                                    (D == 2 && a.cfl_now && stage == 1 && dir == 1))) {
        // (a periodic run: FP32-flux 3D, the stage-1 x faces advance the clock;
        // 2D after the x faces took cfl_m, the stage-1 y faces without the
        // fused y march)
        return;
    }
    const int ix = a.ix, jx = a.jx, kx = a.kx;
    const int n[3] = {ix, jx, kx};
    const int lo_n = stage == 1 ? 1 : 2, hi_n = stage == 1 ? n[dir] - 1 : n[dir] - 2;
    const int lo_t = stage == 1 ? 2 : 3;
    int lo[3], cnt[3];
    for (int d = 0; d < 3; d++) {
        if (d >= D) { lo[d] = 1; cnt[d] = 1; continue; }
        if (d == dir) { lo[d] = lo_n; cnt[d] = hi_n - lo_n + 1; }
        else { lo[d] = lo_t; cnt[d] = n[d] + 1 - 2 * lo_t + 1; }
        if (cnt[d] < 0) cnt[d] = 0;
    }
    const size_t total = (size_t)cnt[0] * cnt[1] * cnt[2];
    const size_t t = (D < 3 ? xcd_first_thread() : (size_t)blockIdx.x * blockDim.x) + threadIdx.x;
    if (CFL) {   // every thread of the block reaches the block minimum
        double cq = kR4_1e20;
        if (t < total) roe_face<D, T, C3, true>(a, dir, stage, lo, cnt, t, cq);
        block_min_to_global(cq, red, &a.clk->dtkey[a.step % 3]);
        return;
    }
    if (t >= total) return;
    double cq;
    roe_face<D, T, C3, false>(a, dir, stage, lo, cnt, t, cq);
}

// 3D y and z faces, marched: thread (i, transverse index, segment) walks the
// faces f = fb..fe of its segment along dir with the four cells f-1..f+2 of
// each face in a sliding register window, one new cell (8 loads) per face
// instead of 32.  Cells outside [1, n] are clamped; they only feed stage-1
// faces, which read cells f and f+1 alone.  Same face range, states and
// flux as roe_flux_kernel<3, T, false>.
template <class T>
__global__ __launch_bounds__(256) void roe_flux_march3(RoeArgs a, int dir, int stage, int H) {
    double dtv;
    // (FP64 periodic run: the stage-1 y faces advance the clock, after the
    // x faces took the step's cfl_m)
    if (!step_dt(a, dtv, sizeof(T) == 8 && a.cfl_now && stage == 1 && dir == 1)) return;
    const int ix = a.ix, jx = a.jx, kx = a.kx;
    const int n[3] = {ix, jx, kx};
    const int lo_n = stage == 1 ? 1 : 2, hi_n = stage == 1 ? n[dir] - 1 : n[dir] - 2;
    const int lo_t = stage == 1 ? 2 : 3;
    const int od = dir == 1 ? 2 : 1;                  // the other transverse direction
    const int ci = ix + 2 - 2 * lo_t, co = n[od] + 2 - 2 * lo_t, nf = hi_n - lo_n + 1;
    if (ci <= 0 || co <= 0 || nf <= 0) return;
    const int nseg = (nf + H - 1) / H;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)ci * co * nseg) return;
    const int i = lo_t + (int)(t % ci);
    const int o = lo_t + (int)((t / ci) % co);
    const int seg = (int)(t / ((size_t)ci * co));
    const int fb = lo_n + seg * H, fe = min(fb + H - 1, hi_n);
    const int nd = n[dir];
    const size_t s = dir == 1 ? (size_t)ix : (size_t)ix * jx;
    const size_t base = dir == 1 ? lin(i, 1, o, ix, jx) : lin(i, o, 1, ix, jx);   // cell 1 along dir
    const double *const *P = stage == 1 ? a.prim : a.half;
    double w0[8], w1[8], w2[8], w3[8];
    auto load_cell = [&](int c, double *d) {
        const int cc = c < 1 ? 1 : (c > nd ? nd : c);
        const size_t pc = base + s * (size_t)(cc - 1);
#pragma unroll
        for (int q = 0; q < 8; q++) d[q] = P[q][pc];
    };
    load_cell(fb - 1, w0); load_cell(fb, w1); load_cell(fb + 1, w2); load_cell(fb + 2, w3);
    const T gm = T(a.gm);
    double *const *F = a.flux[dir];
    const int vn = S_X + dir, v1 = S_X + (dir + 1) % 3, v2 = S_X + (dir + 2) % 3;
    const int bn = S_BX + dir, b1 = S_BX + (dir + 1) % 3, b2 = S_BX + (dir + 2) % 3;
    for (int f = fb;; f++) {
        T L[8], R[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            if (stage == 1) {
                L[q] = T(w1[q]);
                R[q] = T(w2[q]);
            } else {               // cans3d/hdroe/tvdminmod.f:15-55
                L[q] = T(w1[q] + 0.5 * flmt(w1[q] - w0[q], w2[q] - w1[q]));
                R[q] = T(w2[q] - 0.5 * flmt(w3[q] - w2[q], w2[q] - w1[q]));
            }
        }
        T out[7];
        roeflux3<T>(gm, L[S_RO], L[S_E], L[vn], L[v1], L[v2], L[bn], L[b1], L[b2], R[S_RO], R[S_E], R[vn], R[v1],
                    R[v2], R[bn], R[b1], R[b2], out);
        const size_t p = base + s * (size_t)(f - 1);
        fput<T>(F[S_RO], p, out[0]); fput<T>(F[S_E], p, out[1]);
        fput<T>(F[vn], p, out[2]); fput<T>(F[v1], p, out[3]); fput<T>(F[v2], p, out[4]);
        fput<T>(F[b1], p, out[5]); fput<T>(F[b2], p, out[6]);
        if (f >= fe) break;
#pragma unroll
        for (int q = 0; q < 8; q++) { w0[q] = w1[q]; w1[q] = w2[q]; w2[q] = w3[q]; }
        load_cell(f + 3, w3);
    }
}

// conserved variables of a cell from the primitives, in the reference's
// expressions (roe_m.f: rx=ro*vx, ee=pr/(gm-1)+0.5*ro*v2+pi8i*b2)
template <int D, bool C3>
CANS_DEV void prim_to_cons(const RoeArgs &a, size_t p, int i, double *U) {
    const double pi8i = 5.0e-1 * (1.0 / (4.0 * kPi));
    const double gm = a.gm;
    const double ro = a.prim[S_RO][p], pr = a.prim[S_E][p], vx = a.prim[S_X][p], vy = a.prim[S_Y][p];
    U[S_RO] = ro;
    U[S_X] = ro * vx;
    U[S_Y] = ro * vy;
    if (D == 3 || C3) {
        const double vz = a.prim[S_Z][p], by = a.prim[S_BY][p], bz = a.prim[S_BZ][p];
        const double bx = D == 1 ? a.bx[i - 1] : a.prim[S_BX][p];
        U[S_Z] = ro * vz;
        const double v2 = vx * vx + vy * vy + vz * vz;
        const double b2 = bx * bx + by * by + bz * bz;
        U[S_E] = pr / (gm - 1.0) + 0.5 * ro * v2 + pi8i * b2;
        U[S_BX] = bx; U[S_BY] = by; U[S_BZ] = bz;
    } else {
        const double bx = D == 1 ? a.bx[i - 1] : a.prim[S_BX][p], by = a.prim[S_BY][p];
        const double v2 = vx * vx + vy * vy;
        const double b2 = bx * bx + by * by;
        U[S_E] = pr / (gm - 1.0) + 0.5 * ro * v2 + pi8i * b2;
        U[S_BX] = bx; U[S_BY] = by;
    }
}

// ------------------------------------------------------------------ predictor
// cells [2,n-1]^D: U + 1/2 dt (F(i-1)-F(i))/dx, sweeps in x, y, z order
// (each sweep a separate rounding, as roe_m.f); skip the face-normal field.
// Cells on the outer layer keep their primitives (the reference leaves the
// half-step there unset; it only reaches ghost cells when margin >= 3).
template <int D, class T, bool C3>
__global__ __launch_bounds__(256) void roe_half_kernel(RoeArgs a) {
    constexpr bool FULL = D == 3 || C3;
    double dt;
    if (!step_dt(a, dt)) return;
    const int ix = a.ix, jx = a.jx, kx = a.kx;
    const size_t N = (size_t)ix * jx * kx;
    const size_t p = (D < 3 ? xcd_first_thread() : (size_t)blockIdx.x * blockDim.x) + threadIdx.x;
    if (p >= N) return;
    const int i = (int)(p % ix) + 1;
    const int j = (int)((p / ix) % jx) + 1;
    const int k = (int)(p / ((size_t)ix * jx)) + 1;
    const int idx[3] = {i, j, k};
    const int n[3] = {ix, jx, kx};
    bool inner = true;
    for (int d = 0; d < D; d++) inner = inner && idx[d] >= 2 && idx[d] <= n[d] - 1;
    constexpr int NS = 8;
    if (!inner) {
        for (int q = 0; q < NS; q++)
            if (a.prim[q]) a.half[q][p] = a.prim[q][p];
        return;
    }
    double U[8];
    prim_to_cons<D, C3>(a, p, i, U);
    const double *dd[3] = {a.dx, a.dy, a.dz};
    for (int d = 0; d < D; d++) {
        const size_t s = d == 0 ? 1 : (d == 1 ? (size_t)ix : (size_t)ix * jx);
        const double dl = dd[d][idx[d] - 1];
        double *const *F = a.flux[d];
        const int skipB = D == 1 ? S_BX : S_BX + d;    // no normal-field update
        for (int q = 0; q < 8; q++) {
            if (q == skipB || !a.prim[q] || (D == 1 && q == S_BX)) continue;
            if (!FULL && (q == S_Z || q == S_BZ)) continue;
            U[q] = U[q] + 0.5 * dt * ((fget<T>(F[q], p - s) - fget<T>(F[q], p)) / dl);
        }
    }
    // half-step primitives (roe_m.f; 1D: b2 with the cell bx, roe_m.f:94-100)
    const double pi8i = 5.0e-1 * (1.0 / (4.0 * kPi));
    const double gm = a.gm;
    const double ro = U[S_RO];
    const double vx = U[S_X] / ro;
    const double vy = U[S_Y] / ro;
    double v2, b2, vz = 0.0;
    if (FULL) {
        vz = U[S_Z] / ro;
        v2 = vx * vx + vy * vy + vz * vz;
        b2 = U[S_BX] * U[S_BX] + U[S_BY] * U[S_BY] + U[S_BZ] * U[S_BZ];
    } else {
        v2 = vx * vx + vy * vy;
        b2 = U[S_BX] * U[S_BX] + U[S_BY] * U[S_BY];
    }
    a.half[S_RO][p] = ro;
    a.half[S_E][p] = (gm - 1.0) * (U[S_E] - 0.5 * ro * v2 - pi8i * b2);
    a.half[S_X][p] = vx;
    a.half[S_Y][p] = vy;
    a.half[S_BY][p] = U[S_BY];
    if (D >= 2) a.half[S_BX][p] = U[S_BX];
    if (FULL) { a.half[S_Z][p] = vz; a.half[S_BZ][p] = U[S_BZ]; }
}

// ------------------------------------------------------------------ corrector
// cells [3,n-2]^D: U + dt (F(i-1)-F(i))/dx per sweep, then (2D) the Az
// source from the half-step state, then primitives -> out arrays.
template <int D, class T, bool C3>
__global__ __launch_bounds__(256) void roe_update_kernel(RoeArgs a) {
    constexpr bool FULL = D == 3 || C3;
    double dt;
    if (!step_dt(a, dt)) return;
    const int ix = a.ix, jx = a.jx, kx = a.kx;
    const int n[3] = {ix, jx, kx};
    int cnt[3];
    for (int d = 0; d < 3; d++) cnt[d] = d < D ? n[d] - 4 : 1;
    if (cnt[0] <= 0 || cnt[1] <= 0 || cnt[2] <= 0) return;
    const size_t t = (D < 3 ? xcd_first_thread() : (size_t)blockIdx.x * blockDim.x) + threadIdx.x;
    if (t >= (size_t)cnt[0] * cnt[1] * cnt[2]) return;
    const int i = 3 + (int)(t % cnt[0]);
    const int j = D >= 2 ? 3 + (int)((t / cnt[0]) % cnt[1]) : 1;
    const int k = D == 3 ? 3 + (int)(t / ((size_t)cnt[0] * cnt[1])) : 1;
    const int idx[3] = {i, j, k};
    const size_t p = lin(i, j, k, ix, jx);
    double U[8];
    prim_to_cons<D, C3>(a, p, i, U);
    const double *dd[3] = {a.dx, a.dy, a.dz};
    for (int d = 0; d < D; d++) {
        const size_t s = d == 0 ? 1 : (d == 1 ? (size_t)ix : (size_t)ix * jx);
        const double dl = dd[d][idx[d] - 1];
        double *const *F = a.flux[d];
        const int skipB = S_BX + d;
        for (int q = 0; q < 8; q++) {
            if (q == skipB || (D == 1 && q == S_BX)) continue;
            if (!FULL && (q == S_Z || q == S_BZ)) continue;
            if (D == 1 && q == S_BX) continue;
            U[q] = U[q] + dt * ((fget<T>(F[q], p - s) - fget<T>(F[q], p)) / dl);
        }
    }
    if (D == 2) {   // roe_m.f:187-193
        const double vxh = a.half[S_X][p], vyh = a.half[S_Y][p], bxh = a.half[S_BX][p], byh = a.half[S_BY][p];
        const double ez = -vxh * byh + vyh * bxh;
        const double saz = -ez;
        a.out_az[p] = a.az[p] + dt * saz;
    }
    const double pi8i = 5.0e-1 * (1.0 / (4.0 * kPi));
    const double gm = a.gm;
    const double ro = U[S_RO];
    const double vx = U[S_X] / ro;
    const double vy = U[S_Y] / ro;
    double v2, b2;
    if (FULL) {
        const double vz = U[S_Z] / ro;
        v2 = vx * vx + vy * vy + vz * vz;
        b2 = U[S_BX] * U[S_BX] + U[S_BY] * U[S_BY] + U[S_BZ] * U[S_BZ];
        a.out[S_Z][p] = vz;
        a.out[S_BZ][p] = U[S_BZ];
    } else {
        const double bx = D == 1 ? a.bx[i - 1] : U[S_BX];
        v2 = vx * vx + vy * vy;
        b2 = bx * bx + U[S_BY] * U[S_BY];
    }
    a.out[S_RO][p] = ro;
    a.out[S_E][p] = (gm - 1.0) * (U[S_E] - 0.5 * ro * v2 - pi8i * b2);
    a.out[S_X][p] = vx;
    a.out[S_Y][p] = vy;
    a.out[S_BY][p] = U[S_BY];
    if (D >= 2) a.out[S_BX][p] = U[S_BX];
}

// ------------------------------------------------- 3D: z faces with the update
// The z sweep of a stage fused with the cell update that follows it (the
// predictor of stage 1, the corrector of stage 2): a thread marches the
// z faces of its column as roe_flux_march3 does and, once it holds the
// faces below and above cell k, updates the cell -- the x and y face
// contributions read from their arrays, the z contribution from registers,
// in the order and with the expressions of roe_half_kernel and
// roe_update_kernel.  The z faces never go to memory.  A segment starts one
// face early (computed twice, by the same operations) so that it holds the
// face below its first cell.
template <class T, int STAGE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && STAGE == 2 ? 2 : 1))) void roe_zmarch3(RoeArgs a, int H) {
    __shared__ double red[4];
    // per thread: the image-minimum dx, dy of its column and its running cfl
    // minimum (in LDS: the march holds about 250 VGPRs already)
    __shared__ double cfs[2][256];
    double dt;
    if (!step_dt(a, dt)) return;
    constexpr int dir = 2;
    const int ix = a.ix, jx = a.jx, kx = a.kx;
    const int lo_n = STAGE == 1 ? 1 : 2, hi_n = STAGE == 1 ? kx - 1 : kx - 2;
    const int lo_t = STAGE == 1 ? 2 : 3;
    const int ci = ix + 2 - 2 * lo_t, co = jx + 2 - 2 * lo_t, nf = hi_n - lo_n + 1;
    if (ci <= 0 || co <= 0 || nf <= 0) return;
    const int nseg = (nf + H - 1) / H;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    // the next step's cfl_m in a periodic run (stage 2 with FP32 fluxes):
    // every thread of the block reaches the block minimum at the end.  The
    // FP64 march keeps the separate cfl pass: with the term compiled in, its
    // loop waits on its loads twice as often (20.0 -> 23.2 ms per 384^3 step).
    constexpr bool kCfl = STAGE == 2 && sizeof(T) == 4;
    const int mg = a.per_margin;
    const bool do_cfl = kCfl && a.clk && mg > 0 && a.safety > 0.0;
    const int tl = threadIdx.x;
    if (kCfl && do_cfl) cfs[1][tl] = kR4_1e20;
    if (t < (size_t)ci * co * nseg) {
        const int i = lo_t + (int)(t % ci);
        const int j = lo_t + (int)((t / ci) % co);
        const int seg = (int)(t / ((size_t)ci * co));
        const int fb = lo_n + seg * H, fe = min(fb + H - 1, hi_n);
        const int fs = fb > lo_n ? fb - 1 : fb;
        const int nd = kx;
        const size_t sx = 1, sy = (size_t)ix, s = (size_t)ix * jx;
        const size_t base = lin(i, j, 1, ix, jx);
        const double *const *P = STAGE == 1 ? a.prim : a.half;
        double w0[8], w1[8], w2[8], w3[8];
        auto load_cell = [&](int c, double *d) {
            const int cc = c < 1 ? 1 : (c > nd ? nd : c);
            const size_t pc = base + s * (size_t)(cc - 1);
#pragma unroll
            for (int q = 0; q < 8; q++) d[q] = P[q][pc];
        };
        load_cell(fs - 1, w0); load_cell(fs, w1); load_cell(fs + 1, w2); load_cell(fs + 2, w3);
        const T gm = T(a.gm);
        constexpr int vn = S_Z, v1 = S_X, v2 = S_Y, bn = S_BZ, b1 = S_BX, b2 = S_BY;
        const double dx = a.dx[i - 1], dy = a.dy[j - 1];
        const double hdt = STAGE == 1 ? 0.5 * dt : dt;
        // interior column [m+1, n-m]^2 of the cfl_m (periodic images: image_min)
        const bool cij = kCfl && do_cfl && i >= mg + 1 && i <= ix - mg && j >= mg + 1 && j <= jx - mg;
        if (kCfl && cij) cfs[0][tl] = fminF(image_min(a.dx, i, ix, mg), image_min(a.dy, j, jx, mg));
        double Fl[8];   // z flux of the face below the cell, U slots (normal field: unused)
#pragma unroll
        for (int q = 0; q < 8; q++) Fl[q] = 0.0;
        const double pi8i = 5.0e-1 * (1.0 / (4.0 * kPi));
        // With FP32 fluxes (registers to spare) the cell update's operands are
        // loaded before the face flux is computed, so that the flux covers
        // their latency; with FP64 fluxes they would spill.
        constexpr bool kPre = sizeof(T) == 4;
        for (int f = fs;; f++) {
            double pw[8], fxl[8], fxh[8], fyl[8], fyh[8];
            if (kPre && f > fs) {
                const size_t p = base + s * (size_t)(f - 1);
                double *const *Fx = a.flux[0];
                double *const *Fy = a.flux[1];
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    pw[q] = a.prim[q][p];
                    if (q != S_BX) { fxl[q] = fget<T>(Fx[q], p - sx); fxh[q] = fget<T>(Fx[q], p); }
                    if (q != S_BY) { fyl[q] = fget<T>(Fy[q], p - sy); fyh[q] = fget<T>(Fy[q], p); }
                }
            }
            T L[8], R[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                if (STAGE == 1) {
                    L[q] = T(w1[q]);
                    R[q] = T(w2[q]);
                } else {               // cans3d/hdroe/tvdminmod.f:15-55
                    L[q] = T(w1[q] + 0.5 * flmt(w1[q] - w0[q], w2[q] - w1[q]));
                    R[q] = T(w2[q] - 0.5 * flmt(w3[q] - w2[q], w2[q] - w1[q]));
                }
            }
            T out[7];
            roeflux3<T>(gm, L[S_RO], L[S_E], L[vn], L[v1], L[v2], L[bn], L[b1], L[b2], R[S_RO], R[S_E], R[vn], R[v1],
                        R[v2], R[bn], R[b1], R[b2], out);
            // the flux as the face arrays hold it (FP32 runs store FP32)
            double Fh[8];
            Fh[S_RO] = double(out[0]); Fh[S_E] = double(out[1]); Fh[vn] = double(out[2]); Fh[v1] = double(out[3]);
            Fh[v2] = double(out[4]); Fh[b1] = double(out[5]); Fh[b2] = double(out[6]); Fh[bn] = 0.0;
            if (f > fs) {
                // cell k = f: faces k-1 (Fl) and k (Fh) along z
                const int k = f;
                const size_t p = base + s * (size_t)(k - 1);
                double U[8];
                const double dz = a.dz[k - 1];
                double *const *Fx = a.flux[0];
                double *const *Fy = a.flux[1];
                if (kPre) {
                    // prim_to_cons<3> on the loaded primitives
                    const double ro = pw[S_RO], pr = pw[S_E], vx = pw[S_X], vy = pw[S_Y], vz = pw[S_Z];
                    const double bx = pw[S_BX], by = pw[S_BY], bz = pw[S_BZ];
                    U[S_RO] = ro;
                    U[S_X] = ro * vx;
                    U[S_Y] = ro * vy;
                    U[S_Z] = ro * vz;
                    const double v2 = vx * vx + vy * vy + vz * vz;
                    const double b2 = bx * bx + by * by + bz * bz;
                    U[S_E] = pr / (a.gm - 1.0) + 0.5 * ro * v2 + pi8i * b2;
                    U[S_BX] = bx; U[S_BY] = by; U[S_BZ] = bz;
                } else {
                    prim_to_cons<3, false>(a, p, i, U);
                }
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    if (q == S_BX) continue;
                    const double lo = kPre ? fxl[q] : fget<T>(Fx[q], p - sx), hi = kPre ? fxh[q] : fget<T>(Fx[q], p);
                    U[q] = U[q] + hdt * ((lo - hi) / dx);
                }
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    if (q == S_BY) continue;
                    const double lo = kPre ? fyl[q] : fget<T>(Fy[q], p - sy), hi = kPre ? fyh[q] : fget<T>(Fy[q], p);
                    U[q] = U[q] + hdt * ((lo - hi) / dy);
                }
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    if (q == S_BZ) continue;
                    U[q] = U[q] + hdt * ((Fl[q] - Fh[q]) / dz);
                }
                const double ro = U[S_RO];
                const double vx = U[S_X] / ro;
                const double vy = U[S_Y] / ro;
                const double vz = U[S_Z] / ro;
                const double v2s = vx * vx + vy * vy + vz * vz;
                const double b2s = U[S_BX] * U[S_BX] + U[S_BY] * U[S_BY] + U[S_BZ] * U[S_BZ];
                double *const *O = STAGE == 1 ? a.half : a.out;
                const double pr = (a.gm - 1.0) * (U[S_E] - 0.5 * ro * v2s - pi8i * b2s);
                O[S_RO][p] = ro;
                O[S_E][p] = pr;
                O[S_X][p] = vx;
                O[S_Y][p] = vy;
                O[S_Z][p] = vz;
                O[S_BX][p] = U[S_BX];
                O[S_BY][p] = U[S_BY];
                O[S_BZ][p] = U[S_BZ];
                if (kCfl && cij && k >= mg + 1 && k <= kx - mg) {
                    // cfl_m of the stored values (v2s, b2s are its sums)
                    const double dtq = cfl3_term(a.safety, a.gm, ro, pr, v2s, b2s,
                                                 fminF(cfs[0][tl], image_min(a.dz, k, kx, mg)));
                    if (dtq < cfs[1][tl]) cfs[1][tl] = dtq;   // NaN never wins
                }
            }
            if (f >= fe) break;
#pragma unroll
            for (int q = 0; q < 8; q++) { Fl[q] = Fh[q]; w0[q] = w1[q]; w1[q] = w2[q]; w2[q] = w3[q]; }
            load_cell(f + 3, w3);
        }
    }
    if (kCfl && do_cfl) block_min_to_global(cfs[1][tl], red, &a.clk->dtkey[(a.step + 1) % 3]);
}

// Half-step state of the outer cell layer of a 3D grid: the primitives
// themselves (the predictor updates only cells [2, n-1]^3).
__global__ __launch_bounds__(256) void roe_half_shell3(RoeArgs a) {
    double dt;
    if (!step_dt(a, dt)) return;
    const int ix = a.ix, jx = a.jx, kx = a.kx;
    const size_t fx = (size_t)jx * kx, fy = (size_t)(ix - 2) * kx, fz = (size_t)(ix - 2) * (jx - 2);
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (t < 2 * fx) {
        const size_t r = t % fx;
        i = t < fx ? 1 : ix;
        j = 1 + (int)(r % jx);
        k = 1 + (int)(r / jx);
    } else if (t < 2 * fx + 2 * fy) {
        const size_t u = t - 2 * fx, r = u % fy;
        j = u < fy ? 1 : jx;
        i = 2 + (int)(r % (ix - 2));
        k = 1 + (int)(r / (ix - 2));
    } else if (t < 2 * fx + 2 * fy + 2 * fz) {
        const size_t u = t - 2 * fx - 2 * fy, r = u % fz;
        k = u < fz ? 1 : kx;
        i = 2 + (int)(r % (ix - 2));
        j = 2 + (int)(r / (ix - 2));
    } else {
        return;
    }
    const size_t p = lin(i, j, k, ix, jx);
#pragma unroll
    for (int q = 0; q < 8; q++) a.half[q][p] = a.prim[q][p];
}

// ------------------------------------------------- 2D: y faces with the update
// The 2D form of roe_zmarch3: a thread marches the y faces of column i and
// applies the stage's update (predictor or corrector, plus the corrector's
// Az source) to each cell between two of them; x faces from their array.
template <class T, int STAGE>
__global__ __launch_bounds__(256) void roe_ymarch2(RoeArgs a, int H) {
    double dt;
    if (!step_dt(a, dt)) return;
    const int ix = a.ix, jx = a.jx;
    const int lo_n = STAGE == 1 ? 1 : 2, hi_n = STAGE == 1 ? jx - 1 : jx - 2;
    const int lo_t = STAGE == 1 ? 2 : 3;
    const int ci = ix + 2 - 2 * lo_t, nf = hi_n - lo_n + 1;
    if (ci <= 0 || nf <= 0) return;
    const int nseg = (nf + H - 1) / H;
    const size_t t = xcd_first_thread() + threadIdx.x;
    if (t >= (size_t)ci * nseg) return;
    const int i = lo_t + (int)(t % ci);
    const int seg = (int)(t / ci);
    const int fb = lo_n + seg * H, fe = min(fb + H - 1, hi_n);
    const int fs = fb > lo_n ? fb - 1 : fb;
    const int nd = jx;
    const size_t s = (size_t)ix;
    const size_t base = lin(i, 1, 1, ix, jx);
    const double *const *P = STAGE == 1 ? a.prim : a.half;
    constexpr int NS = 6;
    constexpr int slots[NS] = {S_RO, S_E, S_X, S_Y, S_BX, S_BY};
    double w0[NS], w1[NS], w2[NS], w3[NS];
    auto load_cell = [&](int c, double *d) {
        const int cc = c < 1 ? 1 : (c > nd ? nd : c);
        const size_t pc = base + s * (size_t)(cc - 1);
#pragma unroll
        for (int m = 0; m < NS; m++) d[m] = P[slots[m]][pc];
    };
    load_cell(fs - 1, w0); load_cell(fs, w1); load_cell(fs + 1, w2); load_cell(fs + 2, w3);
    const T gm = T(a.gm);
    const double dx = a.dx[i - 1];
    const double hdt = STAGE == 1 ? 0.5 * dt : dt;
    const double pi8i = 5.0e-1 * (1.0 / (4.0 * kPi));
    double Fl[8];
#pragma unroll
    for (int q = 0; q < 8; q++) Fl[q] = 0.0;
    for (int f = fs;; f++) {
        T L[8], R[8];
#pragma unroll
        for (int q = 0; q < 8; q++) { L[q] = T(0); R[q] = T(0); }
#pragma unroll
        for (int m = 0; m < NS; m++) {
            const int q = slots[m];
            if (STAGE == 1) {
                L[q] = T(w1[m]);
                R[q] = T(w2[m]);
            } else {               // cans2d/hdroe/tvdminmod.f
                const double d0 = w0[m], d1 = w1[m], d2 = w2[m], d3 = w3[m];
                L[q] = T(d1 + 0.5 * flmt(d2 - d1, d1 - d0));
                R[q] = T(d2 - 0.5 * flmt(d2 - d1, d3 - d2));
            }
        }
        // 2D y: (vy, vx, by, bx) -> ro ee ry rx bx
        T o[5];
        roeflux12<T>(gm, L[S_RO], L[S_E], L[S_Y], L[S_X], L[S_BY], L[S_BX], R[S_RO], R[S_E], R[S_Y], R[S_X], R[S_BY],
                     R[S_BX], o);
        double Fh[8];
#pragma unroll
        for (int q = 0; q < 8; q++) Fh[q] = 0.0;
        Fh[S_RO] = double(o[0]); Fh[S_E] = double(o[1]); Fh[S_Y] = double(o[2]); Fh[S_X] = double(o[3]);
        Fh[S_BX] = double(o[4]);
        if (f > fs) {
            const int j = f;
            const size_t p = base + s * (size_t)(j - 1);
            double U[8];
            prim_to_cons<2, false>(a, p, i, U);
            const double dy = a.dy[j - 1];
            double *const *Fx = a.flux[0];
            // x then y, as roe_half_kernel / roe_update_kernel (2D: no z
            // components; no update of the face-normal field)
            const int qs[5] = {S_RO, S_E, S_X, S_Y, S_BY};
#pragma unroll
            for (int m = 0; m < 5; m++) {
                const int q = qs[m];
                U[q] = U[q] + hdt * ((fget<T>(Fx[q], p - 1) - fget<T>(Fx[q], p)) / dx);
            }
            const int qy[5] = {S_RO, S_E, S_X, S_Y, S_BX};
#pragma unroll
            for (int m = 0; m < 5; m++) {
                const int q = qy[m];
                U[q] = U[q] + hdt * ((Fl[q] - Fh[q]) / dy);
            }
            if (STAGE == 2) {   // roe_m.f:187-193
                const double vxh = a.half[S_X][p], vyh = a.half[S_Y][p], bxh = a.half[S_BX][p], byh = a.half[S_BY][p];
                const double ez = -vxh * byh + vyh * bxh;
                const double saz = -ez;
                a.out_az[p] = a.az[p] + dt * saz;
            }
            const double ro = U[S_RO];
            const double vx = U[S_X] / ro;
            const double vy = U[S_Y] / ro;
            const double v2 = vx * vx + vy * vy;
            const double b2 = U[S_BX] * U[S_BX] + U[S_BY] * U[S_BY];
            double *const *O = STAGE == 1 ? a.half : a.out;
            O[S_RO][p] = ro;
            O[S_E][p] = (a.gm - 1.0) * (U[S_E] - 0.5 * ro * v2 - pi8i * b2);
            O[S_X][p] = vx;
            O[S_Y][p] = vy;
            O[S_BY][p] = U[S_BY];
            O[S_BX][p] = U[S_BX];
        }
        if (f >= fe) break;
#pragma unroll
        for (int q = 0; q < 8; q++) Fl[q] = Fh[q];
#pragma unroll
        for (int m = 0; m < NS; m++) { w0[m] = w1[m]; w1[m] = w2[m]; w2[m] = w3[m]; }
        load_cell(f + 3, w3);
    }
}

// Half-step state of the outer cell ring of a 2D grid: its primitives.
__global__ __launch_bounds__(256) void roe_half_ring2(RoeArgs a) {
    double dt;
    // (a periodic run after the x faces took cfl_m: this kernel advances the clock)
    if (!step_dt(a, dt, a.cfl_now != 0)) return;
    const int ix = a.ix, jx = a.jx;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j;
    if (t < 2 * (size_t)ix) {
        i = 1 + (int)(t % ix);
        j = t < (size_t)ix ? 1 : jx;
    } else if (t < 2 * (size_t)ix + 2 * (size_t)(jx - 2)) {
        const size_t u = t - 2 * (size_t)ix;
        j = 2 + (int)(u % (jx - 2));
        i = u < (size_t)(jx - 2) ? 1 : ix;
    } else {
        return;
    }
    const size_t p = lin(i, j, 1, ix, jx);
#pragma unroll
    for (int q = 0; q < 8; q++)
        if (a.prim[q]) a.half[q][p] = a.prim[q][p];
}

// 1D and 2D kernels take their blocks in the XCD-contiguous order
// (xcd_first_thread); measured at 1024^2 the 2D engine drops from 294 to
// 266 us.  In 3D the linear order is faster (Roe 24.1 against 25.5 ms at
// 384^3), so the 3D launches keep it.
template <int D> unsigned roe_grid(size_t nthreads) {
    return D < 3 ? xcd_grid(nthreads) : (unsigned)((nthreads + 255) / 256);
}

template <int D, class T, bool C3>
hipError_t launch_roe_t(const RoeArgs &a, hipStream_t s) {
    const int n[3] = {a.ix, a.jx, a.kx};
    const size_t N = (size_t)a.ix * a.jx * a.kx;
    // 2D: y faces fused with the cell updates (roe_ymarch2) on grids with
    // at least 512K stage-1 y faces: segments of >= 8 faces (one in 8 or
    // fewer computed twice) still give >= 1024 waves there.  Smaller grids
    // (and roe_m3) take the separate passes, a thread per face: at 512^2 the
    // march, cut into 3-face segments to fill the chip, was 6 % slower.
    const bool fy2 = D == 2 && !C3 && n[0] >= 3 && n[1] >= 3 &&
                     (size_t)(n[0] - 2) * (size_t)(n[1] - 1) >= ((size_t)1 << 19);
    // 3D: z faces fused with the cell updates (roe_zmarch3) on large grids
    const bool fz3 = D == 3 && roe3d_fused_z(n[0], n[1], n[2]);
    // a periodic run's cfl_m: 3D FP32 fluxes in the fused z marches; 3D FP64
    // and 2D in the stage-1 x faces (cfl_now)
    if (a.per_margin > 0 && !((D == 3 && !C3 && (sizeof(T) == 8 || fz3)) || (D == 2 && !C3)))
        return hipErrorInvalidValue;
    const bool xcfl = !C3 && a.cfl_now && a.per_margin > 0 && ((D == 3 && sizeof(T) == 8) || D == 2);
    for (int stage = 1; stage <= 2; stage++) {
        for (int dir = 0; dir < D; dir++) {
            size_t total = 1;
            for (int d = 0; d < D; d++) {
                long c;
                if (d == dir) c = stage == 1 ? n[d] - 1 : n[d] - 3;
                else c = stage == 1 ? n[d] - 2 : n[d] - 4;
                total *= (size_t)(c > 0 ? c : 0);
            }
            if (total == 0) continue;
            if (fz3 && dir == 2) {
                // z faces fused with the stage's cell update (roe_zmarch3)
                const int nf = stage == 1 ? n[2] - 1 : n[2] - 3;
                const size_t cols = (size_t)(stage == 1 ? n[0] - 2 : n[0] - 4) *
                                    (size_t)(stage == 1 ? n[1] - 2 : n[1] - 4);
                int nseg = (int)(((unsigned long long)march_waves() * 64 + cols - 1) / cols);
                nseg = nseg < 1 ? 1 : (nseg > nf ? nf : nseg);
                int H = (nf + nseg - 1) / nseg;
                H = H < 8 ? 8 : H;
                nseg = (nf + H - 1) / H;
                const size_t threads = cols * (size_t)nseg;
                if (stage == 1) {
                    const size_t shell = 2 * ((size_t)n[1] * n[2] + (size_t)(n[0] - 2) * n[2] +
                                              (size_t)(n[0] - 2) * (n[1] - 2));
                    hipLaunchKernelGGL(roe_half_shell3, dim3((unsigned)((shell + 255) / 256)), dim3(256), 0, s, a);
                    hipLaunchKernelGGL((roe_zmarch3<T, 1>), dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s,
                                       a, H);
                } else {
                    hipLaunchKernelGGL((roe_zmarch3<T, 2>), dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s,
                                       a, H);
                }
                continue;
            }
            if (D == 3 && dir >= 1) {   // y faces (and z faces without the fusion)
                // y and z faces marched (roe_flux_march3: 384^3 FP64 step 26.4 ->
                // 23.7 ms, FP32-flux 19.5 -> 17.7 ms); segments along dir so that
                // the launch has >= march_waves() waves
                const int nf = stage == 1 ? n[dir] - 1 : n[dir] - 3;     // faces along dir
                const int od = dir == 1 ? 2 : 1;
                const size_t cols = (size_t)(stage == 1 ? n[0] - 2 : n[0] - 4) *
                                    (size_t)(stage == 1 ? n[od] - 2 : n[od] - 4);
                int nseg = (int)(((unsigned long long)march_waves() * 64 + cols - 1) / cols);
                nseg = nseg < 1 ? 1 : (nseg > nf ? nf : nseg);
                const int H = (nf + nseg - 1) / nseg;
                nseg = (nf + H - 1) / H;
                const size_t threads = cols * (size_t)nseg;
                hipLaunchKernelGGL((roe_flux_march3<T>), dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, a,
                                   dir, stage, H);
                continue;
            }
            if (fy2 && dir == 1) {
                // y faces fused with the stage's cell update (roe_ymarch2):
                // segments of about 9 faces (one face in 9 computed twice),
                // about 2048 waves at 1024^2
                const int nf = stage == 1 ? n[1] - 1 : n[1] - 3;
                const size_t cols = (size_t)(stage == 1 ? n[0] - 2 : n[0] - 4);
                int nseg = (int)((2048ull * 64 + cols - 1) / cols);
                nseg = nseg < 1 ? 1 : (nseg > nf ? nf : nseg);
                int H = (nf + nseg - 1) / nseg;
                H = H < 8 ? 8 : H;
                nseg = (nf + H - 1) / H;
                const size_t threads = cols * (size_t)nseg;
                if (stage == 1) {
                    const size_t ring = 2 * (size_t)n[0] + 2 * (size_t)(n[1] - 2);
                    hipLaunchKernelGGL(roe_half_ring2, dim3((unsigned)((ring + 255) / 256)), dim3(256), 0, s, a);
                    hipLaunchKernelGGL((roe_ymarch2<T, 1>), dim3(xcd_grid(threads)), dim3(256), 0, s, a, H);
                } else {
                    hipLaunchKernelGGL((roe_ymarch2<T, 2>), dim3(xcd_grid(threads)), dim3(256), 0, s, a, H);
                }
                continue;
            }
            if constexpr (D >= 2 && !C3) {
                if (xcfl && stage == 1 && dir == 0) {
                    hipLaunchKernelGGL((roe_flux_kernel<D, T, C3, true>), dim3(roe_grid<D>(total)), dim3(256), 0, s,
                                       a, dir, stage);
                    continue;
                }
            }
            hipLaunchKernelGGL((roe_flux_kernel<D, T, C3>), dim3(roe_grid<D>(total)), dim3(256), 0, s, a, dir, stage);
        }
        if (stage == 1 && (D == 1 || (D == 2 && !fy2) || (D == 3 && !fz3)))
            hipLaunchKernelGGL((roe_half_kernel<D, T, C3>), dim3(roe_grid<D>(N)), dim3(256), 0, s, a);
    }
    size_t cells = 1;
    for (int d = 0; d < D; d++) cells *= (size_t)(n[d] - 4 > 0 ? n[d] - 4 : 0);
    if (cells && (D == 1 || (D == 2 && !fy2) || (D == 3 && !fz3)))
        hipLaunchKernelGGL((roe_update_kernel<D, T, C3>), dim3(roe_grid<D>(cells)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace

// 3D: z faces fused with the cell updates (roe_zmarch3) where the march's
// segments come out at >= 8 faces for the launch's wave target (about 256^3
// and up at the default); on smaller grids the segments would be a face or
// two long and the fused march would compute most z faces twice, so they
// take the separate passes.
bool roe3d_fused_z(int ix, int jx, int kx) {
    if (ix < 3 || jx < 3 || kx < 3) return false;
    const size_t cols = (size_t)(ix - 2) * (size_t)(jx - 2);
    const int nf = kx - 1;
    long nseg = (long)(((unsigned long long)march_waves() * 64 + cols - 1) / cols);
    nseg = nseg < 1 ? 1 : (nseg > nf ? nf : nseg);
    return (nf + nseg - 1) / nseg >= 8;
}

hipError_t launch_roe2d_half(const RoeArgs &a, hipStream_t s) {
    for (int dir = 0; dir < 2; dir++) {
        const int n[2] = {a.ix, a.jx};
        const size_t total = (size_t)(n[dir] - 1) * (size_t)(n[1 - dir] - 2);
        hipLaunchKernelGGL((roe_flux_kernel<2, double, false>), dim3(roe_grid<2>(total)), dim3(256), 0, s, a, dir, 1);
    }
    hipLaunchKernelGGL((roe_half_kernel<2, double, false>), dim3(roe_grid<2>((size_t)a.ix * a.jx)), dim3(256), 0, s,
                       a);
    return hipGetLastError();
}

hipError_t launch_roe(const RoeArgs &a, int dim, bool mixed, hipStream_t s, bool comp3) {
    if (dim == 1 && comp3) return mixed ? launch_roe_t<1, float, true>(a, s) : launch_roe_t<1, double, true>(a, s);
    if (dim == 2 && comp3) return mixed ? launch_roe_t<2, float, true>(a, s) : launch_roe_t<2, double, true>(a, s);
    if (dim == 1) return mixed ? launch_roe_t<1, float, false>(a, s) : launch_roe_t<1, double, false>(a, s);
    if (dim == 2) return mixed ? launch_roe_t<2, float, false>(a, s) : launch_roe_t<2, double, false>(a, s);
    return mixed ? launch_roe_t<3, float, false>(a, s) : launch_roe_t<3, double, false>(a, s);
}

}  // namespace cans
