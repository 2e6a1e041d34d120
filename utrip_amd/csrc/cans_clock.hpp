// cans_clock.hpp -- device side of the resident time-loop clock.
//
// The reference loop body (e.g. cans2d/md_mhdshktb/hdmlw/main.f:150-206) is
//     ns = ns + 1; cfl_m -> dt; if (merr != 0) exit; t = t + dt;
//     engine(dt); bnd; if (.not.(ns < nstop .and. t < tend)) exit
// clock_enter() evaluates that control flow for step n on the device.
#pragma once
#include "cans_common.hpp"
#include "cans_kernels.hpp"

namespace cans {

// Every block calls this; only the block leader (leader == true) writes the
// next clock slot.  Returns whether the update of step n runs and its dt.
CANS_DEV StepCtl clock_enter(Clock *c, int n, bool leader) {
    const int s = n & 1, d = s ^ 1;
    const bool live = !c->stopped[s] && c->ns[s] < c->nstop && c->t[s] < c->tend;
    const double dt = dunkey(c->dtkey[n % 3]);
    const bool bad = live && (dt < c->dtmin);
    const bool run = live && !bad;
    if (leader) {
        c->ns[d] = live ? c->ns[s] + 1 : c->ns[s];
        c->nexec[d] = run ? c->nexec[s] + 1 : c->nexec[s];
        c->t[d] = run ? c->t[s] + dt : c->t[s];
        c->dt_last[d] = run ? dt : c->dt_last[s];
        c->merr[d] = bad ? 9001 : c->merr[s];
        c->stopped[d] = run ? 0 : 1;
        c->dtkey[(n + 2) % 3] = dkey(kR4_1e20);
    }
    return StepCtl{run, dt};
}

// For kernels that run after the clock kernel of step n: its decision.
CANS_DEV StepCtl clock_peek(const Clock *c, int n) {
    const int d = (n + 1) & 1;
    return StepCtl{c->stopped[d] == 0, c->dt_last[d]};
}

// Step n's run decision and dt, for the engine kernel of the step: advance
// the clock (leader block) unless a clock kernel of the step already did.
CANS_DEV StepCtl clock_use(Clock *c, int n, bool peek, bool leader) {
    return peek ? clock_peek(c, n) : clock_enter(c, n, leader);
}

// Read-only view of the same decision, for kernels of step n other than the
// one that carries the clock (the separate cfl pass, boundary passes).
CANS_DEV bool clock_live(const Clock *c, int n) {
    const int s = n & 1;
    return !c->stopped[s] && c->ns[s] < c->nstop && c->t[s] < c->tend;
}

// ---- the next step's cfl_m fused into the kernel that writes a new state
// In a periodic box the ghost fill that follows the engine (bdperx, bdpery,
// bdperz, in that order) makes every ghost cell a copy of an interior cell,
// its image (i +- (n - 2m) along each axis with a ghost index).  cfl_m over
// every cell of the new state is then the minimum over the interior cells
// [m+1, n-m] of the terms safety*(d/sq), sq from the cell's state and d the
// smallest spacing among the cell's images: m -> safety*(m/sq) is monotone
// in IEEE arithmetic for sq >= 0 (a NaN sq gives NaN terms, which never win
// the reference's `if (dtq < dt)` either way), so that term is bit for bit
// the smallest of the images' terms.

// smallest of d(i) and d at the periodic images of cell i (1-based; n cells,
// margin m, n - 2m >= m so that a cell has at most one image on each side)
CANS_DEV double image_min(const double *d, int i, int n, int m) {
    double g = d[i - 1];
    if (i >= n - 2 * m + 1) g = fminF(g, d[i - (n - 2 * m) - 1]);
    if (i <= 2 * m) g = fminF(g, d[i + (n - 2 * m) - 1]);
    return g;
}

// cans3d/common/cfl_m.f:45-54 for one cell of the new state: ro, pr the
// stored values, v2 = vx*vx + vy*vy + vz*vz and b2 = bx*bx + by*by + bz*bz
// summed in that order, dg = min(dx, dy, dz) over the cell's images;
// onep4 = 0.25/pai with pai the REAL*4 acos(-1.0)
CANS_DEV double cfl3_term(double safety, double gm, double ro, double pr, double v2, double b2, double dg) {
    constexpr double onep4 = 0.25 / kPiF;
    const double ca2 = b2 * onep4 / ro;
    const double cs2 = gm * pr / ro;
    return safety * (dg / __builtin_sqrt(v2 + cs2 + ca2));
}

// cans2d/common/cfl_m.f:46-55, the same with pi4i = 1/pai/4, pai the REAL*8
// acos(-1.0d0), v2 = vx*vx + vy*vy, b2 = bx*bx + by*by, dg = min(dx, dy)
CANS_DEV double cfl2_term(double safety, double gm, double ro, double pr, double v2, double b2, double dg) {
    constexpr double pi4i = 1.0 / kPi / 4.0;
    const double ca2 = b2 * pi4i / ro;
    const double cs2 = gm * pr / ro;
    return safety * (dg / __builtin_sqrt(v2 + cs2 + ca2));
}

}  // namespace cans
