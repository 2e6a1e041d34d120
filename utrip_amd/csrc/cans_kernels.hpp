// cans_kernels.hpp -- launch interface between the C-ABI layer and the HIP
// kernels.  Internal to the library; the public surface is include/cans_gpu.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cans {

// Device-side clock of a resident time loop.  It holds the scalars a CANS
// main program keeps on the host (t, ns, merr) so that a run of steps needs
// no host round trip.  Step n (an index the host passes to every kernel of
// that step) reads the slot n&1 of every scalar and its leader writes slot
// (n+1)&1, so no block ever reads a value another block of the same launch
// writes.  dt travels as order-preserving keys in three rotating slots:
// step n uses dtkey[n%3], min-reduces the next dt into dtkey[(n+1)%3] and
// re-arms dtkey[(n+2)%3] to 1.e20 (the reference's REAL*4 initial value).
struct Clock {
    unsigned long long dtkey[3];
    double t[2];
    double dt_last[2];
    double tend;
    double dtmin;
    int ns[2];       // loop entries, as the reference's ns
    int nexec[2];    // steps whose update ran
    int stopped[2];
    int merr[2];
    int nstop;
    int pad;
};

struct StepCtl {
    bool run;   // this step updates the state
    double dt;
};

// Metric coefficients precomputed once per grid axis; bitwise the same as
// the reference computes them at the top of every engine call.
struct Metric {
    double *di;   // 1.0/d(i)
    double *dim;  // 1.0/dm(i)
    double *u0;   // 0.5*dm(i)/d(i)     (i = 2..n-1)
    double *u1;   // 0.5*dm(i-1)/d(i)   (i = 2..n-1)
};

struct Mlw2dArgs {
    const double *ro, *pr, *vx, *vy, *bx, *by, *az;
    double *ro_o, *pr_o, *vx_o, *vy_o, *bx_o, *by_o, *az_o;
    const double *dx, *dxm, *dy, *dym;
    Metric mx, my;
    double qav, gm, dt;
    Clock *clk;        // non-null: dt and the run flag come from the clock
    int step;          // step index n for the clock
    int ix, jx;
    int per_margin;    // > 0: also write the periodic ghost images (width)
    double safety;     // > 0 with clk: fused cfl_m of the new state
    int peek;          // 1: a clock kernel of this step already ran (read, do not advance)
    double *dt_log;    // non-null: dt_log[step] = the dt this step used
};

struct Mlw1dArgs {
    const double *ro, *pr, *vx, *vy, *by;
    double *ro_o, *pr_o, *vx_o, *vy_o, *by_o;
    const double *bx, *bxm, *dx, *dxm;
    Metric mx;
    double qav, gm, dt;
    Clock *clk;
    int step;
    int ix;
    int peek;          // 1: a clock kernel of this step already ran (read, do not advance)
};

// Boundary kinds (the CANS bc/ routines)
enum BndKind : int {
    BD_FRE = 0,  // bdfre*  zero gradient
    BD_SPP = 1,  // bdspp*  mirror about the cell face, sign kept
    BD_SPN = 2,  // bdspn*  mirror about the cell face, sign flipped
    BD_SMP = 3,  // bdsmp*  mirror about the cell centre, sign kept
    BD_SMN = 4,  // bdsmn*  mirror about the cell centre, sign flipped, wall 0
    BD_CNS = 5,  // bdcns*  constant q0
    BD_FRD = 6,  // bdfrd*  linear extrapolation with dxm
    BD_PER = 7,  // bdper*  periodic, both sides (mbnd = margin0)
    BD_NONE = -1,
};

constexpr int kMaxBndFields = 16;
struct BndArgs {
    double *q[kMaxBndFields];
    double q0[kMaxBndFields];
    int kind[kMaxBndFields];
    int nf;
    int dir;          // 0 x, 1 y, 2 z
    int mbnd;         // side (0 low, 1 high); for BD_PER: margin0
    int margin;       // for BD_PER: margin1
    const double *dm; // dxm/dym/dzm, for BD_FRD
    int ix, jx, kx;
};

struct CflArgs {
    const double *ro, *pr, *vx, *vy, *vz, *bx, *by, *bz;
    const double *dx, *dy, *dz;
    double safety, gm;
    int ix, jx, kx;
    int dim;                    // 1, 2 or 3 (selects the reference's formula)
    int comp3;                  // 1: cfl_m3 of 1D/2D (adds vz, bz)
    const double *xr, *trig;    // spherical cfl_m_s: radii x (2D, 3D) and sin(y) (3D, trig[0..jx)); NULL: cfl_m
    unsigned long long *key;    // order-preserving key of the min (pre-armed)
    Clock *clk;                 // non-null: reduce into clk->dtkey[step%3] if the step runs
    int step;
};

// Batched 1D tubes (packed: tube k occupies cells off[k] .. off[k]+ix[k]-1
// of every array).
struct Batch1dArgs {
    int ntube;
    long long ncells;            // total packed cells
    const long long *off;        // device, per tube
    const int *ix;               // device, per tube (nx + 2*margin)
    const int *order;            // device queue order (NULL: 0..ntube-1)
    double *f[5];                // ro pr vx vy by (state, in place)
    double *g[5];                // scratch set of the same size
    const double *bx, *bxm, *dx, *dxm;
    double *di, *dim;            // 1/dx, 1/dxm (filled by the launch)
    double gm, qav, safety, dtmin, tend, t0;
    int nstop, margin;
    const int *nstop_tube;       // device, per tube (NULL: nstop for every tube)
    double *t_out;               // device, per tube
    int *ns_out, *merr_out;      // device, per tube
    unsigned *next;              // queue counter (zeroed by the launch)
};
hipError_t launch_batch1d(const Batch1dArgs &a, int mixed, int block_threads, int nblocks, hipStream_t s);

// Batched periodic 2D runs (MLW): run k's fields occupy cells off[k] ..
// off[k]+ix[k]*jx[k]-1 of every field array (column-major (ix,jx)); its x
// grid arrays (dx, dxm and the x metric) entries offx[k] .. offx[k]+ix[k]-1,
// its y arrays offy[k] .. offy[k]+jx[k]-1.
struct Batch2dArgs {
    int nrun;
    const long long *off, *offx, *offy;   // device, per run
    const int *ix, *jx;                   // device, per run (ghosts included)
    const int *order;                     // device queue order (NULL: 0..nrun-1)
    double *f[7];                         // ro pr vx vy bx by az (state, in place)
    double *g[7];                         // scratch set of the same size
    const double *dx, *dxm, *dy, *dym;    // packed grid arrays
    double *mx[4], *my[4];                // packed metric (di, dim, u0, u1), filled by the launch
    const Mlw2dArgs *__restrict__ runs;   // per run k: runs[2k] reads f writes g, runs[2k+1] the reverse
    double gm, qav, safety, dtmin, tend;
    int nstop, margin;                    // margin: the periodic ghost width
    double *t_out;                        // device, per run
    int *ns_out, *merr_out;               // device, per run
    unsigned *next;                       // queue counter (zeroed by the launch)
};
hipError_t launch_batch2d(const Batch2dArgs &b, int block_threads, int nblocks, hipStream_t s);

// mlw_m3 (cans{1,2}d/hdmlw/mlw_m3.f): 1D fields ro pr vx vy vz by bz (bx,
// bxm constant); 2D fields ro pr vx vy vz bx by bz az.  in -> out on the
// interior cells (2..n-1).
struct Mlw3Args {
    const double *in[9];
    double *out[9];
    const double *bx, *bxm;          // 1D
    const double *x, *xm;            // cylindrical (mlw_m_c): cell and face radii
    const double *dxm, *dym;
    Metric mx, my;
    double qav, gm, dt;
    Clock *clk;                      // non-null: dt and run flag of step `step` (after its clock kernel)
    int step;
    int ix, jx;
    // 2D mlw_m3 in a periodic box (per_margin = m > 0): the kernel advances
    // the clock itself and takes the next step's cfl_m3 over the interior
    // cells it updates (image_min, cfl2_term; dx, dy the cell widths)
    const double *dx, *dy;
    double safety;
    int per_margin;
};
hipError_t launch_mlw_m3(const Mlw3Args &a, int dim, hipStream_t s);
// mlw_m_c (cans2d/hdmlw/mlw_m_c.f), fields ro pr vx vz bx bz ay on the (x, z)
// plane with the radii a.x, a.xm; my is the z metric
hipError_t launch_mlw_m_c(const Mlw3Args &a, hipStream_t s);

// 2D spherical mlw_m_s (cans2d/hdmlw/mlw_m_s.f): fields ro pr vx vy bx by az,
// SoA scratch of kMlw2sCellFields per cell and kMlw2sCornerFields per corner;
// trig: 6 tables of jx (sin, cos, tan of y, then of ym).
constexpr int kMlw2sCellFields = 19;
constexpr int kMlw2sCornerFields = 13;
struct Mlw2sArgs {
    const double *prim[7];
    double *out[7];                  // interior cells
    double *cell, *corner;
    const double *x, *xm, *trig;
    const double *dxm, *dym;
    Metric mx, my;
    double qav, gm, dt;
    Clock *clk;                      // non-null: dt and run flag of step `step` (after its clock kernel)
    int step;
    int ix, jx;
};
hipError_t launch_mlw2d_s(const Mlw2sArgs &a, hipStream_t s);

// UTRIP spicule (cans1d/md_spicule/hdmlw): geometry arrays of mlw_m_bg, in
// this order.
enum { SG_BX, SG_BXM, SG_GX, SG_GXM, SG_SC, SG_DSC, SG_SCM, SG_DSCM, SG_RR, SG_RRM, SG_DRR, SG_DRRM, SG_DX,
       SG_DXM, SG_N };

// One mlw_m_bg call: in -> out (cells 2..ix-1); met = 1/dx, 1/dxm, ux0, ux1
// as filled by launch_spicule_prep.
struct MlwBgArgs {
    const double *in[5];
    double *out[5];
    const double *geo[SG_N];
    const double *met[4];
    double dt, qav, gm;
    int ix;
};
hipError_t launch_mlw1d_bg(const MlwBgArgs &a, hipStream_t s);

// Many realizations of the spicule time loop (main.f:169-227), each run by
// one workgroup of a persistent kernel.  Run r's fields sit at r*ix of f[];
// its geometry at r*gstride of geo[], met[] and trq (gstride 0: shared).
struct SpiculeArgs {
    int nrun, ix, margin;
    long long gstride;
    double *f[5];                   // ro pr vx vy by (in place)
    double *g[5];                   // scratch of the same size
    const double *geo[SG_N];
    const double *met[4];           // filled by launch_spicule_prep
    const double *trq;              // torque profile (trqalfven.f:23-28)
    const long long *pw;            // 16807^(i+1) mod (2^31-1), i < ix (prep)
    double gm, qav, safety, dtmin, floor, tend, dtout;
    int nstop;
    double *t;                      // per run: in t, out t
    int *ns, *merr, *seed;          // per run: in/out (ns: loop entries so far)
    double *snap, *snap_t;          // NULL, or [run][nsnap_max][5][ix] and [run][nsnap_max]
    int nsnap_max;
    int *nsnap;                     // per run: snapshots written by the launch
    unsigned *next;                 // queue counter (zeroed by the prep kernel)
};
hipError_t launch_spicule(const SpiculeArgs &a, int block_threads, int nblocks, hipStream_t s);
// metric of nset geometry sets (stride gstride), the draw multipliers, and
// the queue counter (may be NULL)
hipError_t launch_spicule_prep(const double *dx, const double *dxm, double *const *met, int ix, int nset,
                               long long gstride, long long *pw, unsigned *next, hipStream_t s);
// trqalfven.f with the Park-Miller stream starting at seed (one call)
hipError_t launch_trq1d(double *vy, double dt, const double *trq, const double *rr, const long long *pw,
                        long long seed, int ix, hipStream_t s);
// chkdav.f on n cells; *count (device) = number of floored cells
hipError_t launch_chkdav1d(double *da, double *vx, double floor_, int n, int *count, hipStream_t s);

// Reverse mode of one fused periodic 2D MLW step (6 fields ro..by).
// B1-B3 of the adjoint of step n.  dtp: the dt step n used; acc: 2 doubles
// of zeroed accumulators (dL/ddt, ties) for the step; pend_dt/pend_acc: dt and
// accumulators of step n+1, whose dt gradient B1 adds (NULL: none);
// acc_zero: accumulators cleared for step n-1 (or NULL).
hipError_t launch_adj2d_step(const double *const *pn, const double *const *pn1, const double *const *lin,
                             double *const *lout, double *ws, const double *dtp, double *acc, const double *pend_dt,
                             const double *pend_acc, double *acc_zero, const double *dx, const double *dxm,
                             const double *dy, const double *dym, Metric mx, Metric my, double qav, double gm,
                             double safety, int ix, int jx, int margin, hipStream_t s);
// the dt gradient of step n added to lout (dL/dP^n) at the cells of P^n that tie
hipError_t launch_adj2d_dt(const double *const *pn, double *const *lout, double *ws, const double *dtp, double *acc,
                           const double *dx, const double *dy, double gm, double safety, int ix, int jx,
                           hipStream_t s);
hipError_t launch_key_to_dt(const unsigned long long *key, double *dt, hipStream_t s);
hipError_t launch_zero(double *p, size_t n, hipStream_t s);
hipError_t launch_periodic_fold(double *const *lam, int ix, int jx, int margin, hipStream_t s);
size_t adj2d_ws_doubles(int ix, int jx);
// reverse mode of one periodic 2D Roe-MUSCL step (margin >= 3); ph: the
// step's half-step primitives (ro pr vx vy bx by) as recorded, or NULL to
// recompute them from pn into ws
hipError_t launch_adj2d_roe_step(const double *const *pn, const double *const *pn1, const double *const *ph,
                                 const double *const *lin, double *const *lout, double *ws, const double *dtp,
                                 double *acc, const double *pend_dt, const double *pend_acc, double *acc_zero,
                                 const double *dx, const double *dy, double gm, double safety, int ix, int jx,
                                 int margin, hipStream_t s);
size_t adj2d_roe_ws_doubles(int ix, int jx);
// One accumulator slot of the adjoint step: dL/ddt in kAdjStripes partial
// sums (blocks add to stripe blockIdx % kAdjStripes, so that same-address
// atomics do not serialize), then the tie count (an unsigned in a double).
constexpr int kAdjStripes = 16;
constexpr int kAdjAccSlot = kAdjStripes + 1;
hipError_t launch_copy6(double *const *dst, double *const *src, size_t n, hipStream_t s);

hipError_t launch_metric(const double *d, const double *dm, int n, Metric m, hipStream_t s);
hipError_t launch_mlw2d(const Mlw2dArgs &a, hipStream_t s);
hipError_t launch_mlw1d(const Mlw1dArgs &a, hipStream_t s);
hipError_t launch_bnd(const BndArgs &a, hipStream_t s);
// every field periodic on every face (margin m along each of the dim axes):
// bdperx, bdpery, bdperz in one pass
hipError_t launch_bnd_periodic(double *const *q, int nf, int ix, int jx, int kx, int dim, int m, hipStream_t s);
hipError_t launch_copy_interior(double *const *dst, const double *const *src, int nf,
                                int ix, int jx, int kx, int lo, hipStream_t s);
hipError_t launch_cfl(const CflArgs &a, hipStream_t s);
// first cell (reference loop order) whose dtq equals dt; loc[0..2] 1-based
hipError_t launch_cfl_locate(const CflArgs &a, double dt, int *loc_dev, hipStream_t s);
hipError_t launch_arm_key(unsigned long long *key, hipStream_t s);  // *key = key(1.e20)
hipError_t launch_clock_reset(Clock *clk, double t0, double tend, double dtmin, int nstop,
                              hipStream_t s);
// a clock-only step (for paths whose update kernel does not carry the clock)
hipError_t launch_clock_step(Clock *clk, int step, hipStream_t s);
hipError_t launch_clock_dt_log(const Clock *clk, int step, double *out, hipStream_t s);

}  // namespace cans
