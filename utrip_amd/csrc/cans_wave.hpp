// cans_wave.hpp -- wavefront-level helpers of the 1D marching kernels
// (64-lane DPP neighbour shifts, wave-uniform values, workgroup min).
#pragma once
#include "cans_common.hpp"

namespace cans {

// Lane l receives lane l+1's value (shl1) or lane l-1's (shr1) through a DPP
// wave shift; the lane with no source receives 0 (bound_ctrl).
CANS_DEV double shl1(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x130, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x130, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}
CANS_DEV double shr1(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x138, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x138, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}
CANS_DEV float shl1(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, true));
}
CANS_DEV float shr1(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, true));
}

// Values every lane of the workgroup holds identically, made uniform for the
// compiler.  Loops that contain barriers must exit through scalar branches:
// with a loop exit the compiler believes divergent, the control-flow
// structurizer may send part of a wave back to a barrier ahead of lanes still
// finishing the previous iteration.
CANS_DEV int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
CANS_DEV double uni(double v) {
    return __hiloint2double(__builtin_amdgcn_readfirstlane(__double2hiint(v)),
                            __builtin_amdgcn_readfirstlane(__double2loint(v)));
}
CANS_DEV long long uni(long long v) {
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v & 0xffffffffll));
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)((unsigned long long)v >> 32));
    return (long long)(((unsigned long long)hi << 32) | lo);
}

// Workgroup min of a double (NaN never wins: callers feed only values that
// passed `v < current`); `red` holds BT/64 doubles.  The result is uniform.
template <int BT>
CANS_DEV double block_min(double v, double *red) {
    constexpr int NW = BT / 64;
    v = wave_min(v);                       // the same value in every lane
    red[threadIdx.x >> 6] = v;             // all lanes store it: no divergent branch
    __syncthreads();
    double m = red[0];
#pragma unroll
    for (int w = 1; w < NW; w++) m = (red[w] < m) ? red[w] : m;
    return uni(m);
}

// one term of the 1D cfl_m (cans1d/common/cfl_m.f:42-52: onero, ca2, cs2,
// dtcfl, safety*dtcfl)
CANS_DEV double dtq1(double ro, double pr, double vx, double by, double bx, double dx, double gm, double safety) {
    constexpr double cfl_pi4i = 0.25 / kPi;
    const double onero = 1.0 / ro;
    const double ca2 = (bx * bx + by * by) * cfl_pi4i * onero;
    const double cs2 = gm * pr * onero;
    const double dtcfl = dx / (__builtin_fabs(vx) + __builtin_sqrt(cs2 + ca2));
    return safety * dtcfl;
}

}  // namespace cans
