// cans_roe.hpp -- argument blocks of the multi-kernel engines (Roe-MUSCL,
// 3D MLW, HLLD).  Internal to the library.
#pragma once
#include "cans_kernels.hpp"

namespace cans {

// Roe + MUSCL (cans{1,2,3}d/hdroe/roe_m.f).  Slots: 0 ro, 1 pr(P)/ee(U),
// 2 vx/rx, 3 vy/ry, 4 vz/rz, 5 bx, 6 by, 7 bz; unused slots are null
// (2D: 4 and 7; 1D: 4, 5 and 7 -- the 1D field bx is `bx`, the face
// field `bxm`).
struct RoeArgs {
    const double *prim[8];   // state at t^n (primitives)
    double *half[8];         // scratch: half-step primitives
    double *flux[3][8];      // scratch: face fluxes per direction (by U slot)
    double *out[8];          // new primitives, cells [3,n-2]^D
    const double *az;        // 2D vector potential (in)
    double *out_az;          // 2D vector potential (out)
    const double *bx, *bxm;  // 1D only
    const double *dx, *dy, *dz;
    double gm, dt;
    Clock *clk;              // non-null: dt and run flag of step `step`
    int step;
    int ix, jx, kx;
    const double *dtp;       // non-null (and clk null): dt read from device memory
    // 3D in a periodic box (per_margin = m > 0).  FP32 fluxes with the fused
    // z march (roe3d_fused_z): the step's first kernel advances the clock and
    // the stage-2 z march takes the next step's cfl_m.  FP64 with cfl_now
    // (not the first step of a call): the stage-1 x faces take this step's
    // cfl_m and the stage-1 y march advances the clock (image_min,
    // cfl3_term).  2D (either precision) with cfl_now: the stage-1 x faces
    // take the step's cfl_m (cfl2_term) and the next kernel (the half-step
    // ring or the stage-1 y faces) advances the clock.
    double safety;
    int per_margin;
    int cfl_now;
};
// comp3: roe_m3 (1D/2D with vz, bz in slots S_Z, S_BZ)
hipError_t launch_roe(const RoeArgs &a, int dim, bool mixed, hipStream_t s, bool comp3 = false);
// whether a 3D Roe step on this grid takes the fused z marches (roe_zmarch3)
bool roe3d_fused_z(int ix, int jx, int kx);
// The first half of a 2D FP64 Roe step: the stage-1 face fluxes of both
// directions (a.flux) and the half-step primitives (a.half); the reverse
// sweep recomputes them from the recorded state.
hipError_t launch_roe2d_half(const RoeArgs &a, hipStream_t s);

// 3D modified Lax-Wendroff (cans3d/hdmlw/mlw_m.f), three kernels with
// global scratch: cell quantities, corner half-step fluxes, update.
struct Mlw3dArgs {
    const double *prim[8];   // ro pr vx vy vz bx by bz (in)
    double *out[8];          // same (out), interior cells
    double *cell;            // scratch: 23 (curvilinear 32) doubles per cell (SoA)
    double *corner;          // scratch: 18 (curvilinear 26) doubles per corner (SoA)
    const double *x, *xm;    // non-null: curvilinear (mlw_m_c / mlw_m_s), cell and face radii
    const double *trig;      // non-null: spherical mlw_m_s; 6 tables of jx: sin, cos, tan of y, then of ym
    const double *dxm, *dym, *dzm;
    Metric mx, my, mz;
    double qav, gm, dt;
    Clock *clk;
    int step;
    int ix, jx, kx;
    // Cartesian marching kernel in a periodic box (per_margin = m > 0): it
    // advances the clock itself (clock_enter; no separate clock kernel for
    // the step) and takes the next step's cfl_m over the new state, the
    // ghost images included through the image-minimum dx, dy, dz
    const double *dx, *dy, *dz;
    double safety;
    int per_margin;
};
hipError_t launch_mlw3d(const Mlw3dArgs &a, hipStream_t s);
constexpr int kMlw3dCellFields = 23;
constexpr int kMlw3dCornerFields = 18;
constexpr int kMlw3dCylCellFields = 32;
constexpr int kMlw3dCylCornerFields = 26;

// HLLD MUSCL-Hancock (cansxx-nosupport/hdhlld, mu units)
struct HlldArgs {
    double *prim[8];         // ro pr vx vy vz bx by bz (in, all cells)
    double *out[8];          // new primitives, all cells (may alias prim)
    double *half[8];         // scratch: half-step primitives
    double *u1[8];           // scratch: conserved state of stage 1
    double *u2[8];           // scratch: conserved state of stage 2
    const double *dx, *dy, *dz;
    double gm, mu, dt;
    int mfdim[3];
    Clock *clk;
    int step;
    int ix, jx, kx;
    // periodic box (per_margin = m > 0; needs the x and z sweeps, mfdim[0]
    // and mfdim[2] = 1): the stage-1 x sweep advances the clock and the
    // stage-2 z sweep takes the next step's cfl_m (image_min, cfl3_term)
    double safety;
    int per_margin;
};
hipError_t launch_hlld(const HlldArgs &a, bool mixed, hipStream_t s);

// Marching face kernels (HLLD y/z sweeps, Roe 3D y/z fluxes) split each
// column into segments so that a launch has at least this many waves
// (default 32768; cans_gpu_set_march_waves changes it).
int march_waves();
// 3D MLW, Cartesian grids: 1 (default) = the z-marching kernel of
// cans_mlw3d_march.hip, 0 = the three kernels over HBM scratch of
// cans_mlw3d.hip (cans_gpu_set_mlw3d_path changes it).
int mlw3d_path();
hipError_t launch_mlw3d_march(const Mlw3dArgs &a, hipStream_t s);

}  // namespace cans
