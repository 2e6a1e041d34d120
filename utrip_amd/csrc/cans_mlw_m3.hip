// cans_mlw_m3.hip -- mlw_m3, the modified Lax-Wendroff engine for MHD with
// all three vector components (cans1d/hdmlw/mlw_m3.f:2-252,
// cans2d/hdmlw/mlw_m3.f:2-377), on gfx950, FP64 bit-identical to the
// reference compiled without FMA contraction.
//
// 1D: a wave updates a segment of 62 cells (lane l holds cell c0-1+l; the
//     face between a cell and its east neighbour belongs to the cell's lane;
//     neighbours travel by DPP shifts), as the other 1D kernels.
// 2D: one workgroup of 256 threads updates a tile of 31 x 7 cells in one
//     pass through LDS: (1) the 33 x 9 cells of the tile and its halo load
//     their state and leave conserved variables, fluxes and velocities in
//     LDS; (2) each thread forms one of the 32 x 8 corner half steps
//     (mlwhalf.f) and its corner fluxes, while the interior threads take the
//     centred half-step term of their cell; (3) the interior threads apply
//     the full step (mlwfull.f), the Az source (mlwsrch.f, mlwsrcf.f), the
//     viscosity (mlwartv.f) and recover the primitives.  Corner values
//     overwrite the cell fluxes in LDS once the last reader is past a
//     barrier, so a tile needs 54.6 KB.
#include "cans_clock.hpp"
#include "cans_common.hpp"
#include "cans_kernels.hpp"
#include "cans_wave.hpp"

namespace cans {
namespace {

constexpr double pi8i = 1.0 / kPi / 8.0;   // mlw_m3.f: pi8i=1./pi/8.
constexpr double pi4i = 1.0 / kPi / 4.0;

// Step n's run decision and dt: after the step's clock kernel, or (2D
// periodic box, per_margin > 0) advanced here, block 0 writing the clock.
CANS_DEV bool m3_step(const Mlw3Args &a, double &dt) {
    if (!a.clk) { dt = a.dt; return true; }
    const StepCtl c = clock_use(a.clk, a.step, a.per_margin == 0, blockIdx.x == 0 && threadIdx.x == 0);
    dt = c.dt;
    return c.run;
}

// ---------------------------------------------------------------- 1D
// fields: ro pr vx vy vz by bz
__global__ __launch_bounds__(256) void mlw1d_m3_kernel(Mlw3Args a) {
    double dt;
    if (!m3_step(a, dt)) return;
    const int lane = threadIdx.x & 63;
    const int seg = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    const int ix = a.ix;
    const int c = 2 + 62 * seg - 1 + lane;                    // this lane's cell, 1-based
    const int p = (c < 1 ? 1 : (c > ix ? ix : c)) - 1;
    const double gm = a.gm;
    const double ro = a.in[0][p], pr = a.in[1][p], vx = a.in[2][p], vy = a.in[3][p], vz = a.in[4][p];
    const double by = a.in[5][p], bz = a.in[6][p], bx = a.bx[p];
    // cells (mlw_m3.f:83-143)
    double u[7], f[7];
    {
        const double vv = vx * vx + vy * vy + vz * vz;
        const double bb = by * by + bz * bz;
        u[0] = ro;
        u[1] = pr / (gm - 1.0) + 0.5 * ro * vv + pi8i * bb;
        u[2] = ro * vx;
        u[3] = ro * vy;
        u[4] = ro * vz;
        u[5] = by;
        u[6] = bz;
        const double ey = -vz * bx + vx * bz;
        const double ez = -vx * by + vy * bx;
        const double ep = pr * gm / (gm - 1.0) + 0.5 * ro * vv;
        f[0] = ro * vx;
        f[1] = ep * vx - by * ez * pi4i + bz * ey * pi4i;
        f[2] = ro * (vx * vx) + pr + bb * pi8i;
        f[3] = ro * vx * vy - bx * by * pi4i;
        f[4] = ro * vx * vz - bx * bz * pi4i;
        f[5] = -ez;
        f[6] = ey;
    }
    // the face to the east: half step and face fluxes (mlw_m3.f:147-207)
    const double dxim = a.mx.dim[p];
    double uE[7], fE[7], ff[7];
    {
        double un[7];
#pragma unroll
        for (int q = 0; q < 7; q++) {
            uE[q] = shl1(u[q]);
            fE[q] = shl1(f[q]);
            const double uh = 0.5 * (uE[q] + u[q]);
            const double dfdx = dxim * (fE[q] - f[q]);
            un[q] = uh - dt * dfdx;
        }
        const double bxm = a.bxm[p];
        const double roh = un[0], eeh = un[1], byh = un[5], bzh = un[6];
        const double vxh = un[2] / roh;
        const double vyh = un[3] / roh;
        const double vzh = un[4] / roh;
        const double vv = vxh * vxh + vyh * vyh + vzh * vzh;
        const double bb = byh * byh + bzh * bzh;
        const double prh = (gm - 1.0) * (eeh - 0.5 * roh * vv - pi8i * bb);
        const double eyh = -vzh * bxm + vxh * bzh;
        const double ezh = -vxh * byh + vyh * bxm;
        const double eph = prh * gm / (gm - 1.0) + 0.5 * roh * vv;
        ff[0] = roh * vxh;
        ff[1] = eph * vxh - byh * ezh * pi4i + bzh * eyh * pi4i;
        ff[2] = roh * (vxh * vxh) + prh + bb * pi8i;
        ff[3] = roh * vxh * vyh - bxm * byh * pi4i;
        ff[4] = roh * vxh * vzh - bxm * bzh * pi4i;
        ff[5] = -ezh;
        ff[6] = eyh;
    }
    // own cell: centred half-step term, full step, viscosity, update
    const double dxi = a.mx.di[p];
    const double zero = 0.0;
    const double qxc = a.qav * a.dxm[p] * fmaxF(zero, __builtin_fabs(shl1(vx) - vx) - kR4_1em4);
    const double qxw = shr1(qxc);
    const double dxim1 = shr1(dxim);
    double nu[7];
#pragma unroll
    for (int q = 0; q < 7; q++) {
        double du = 0.0;
        du = du - 0.5 * dt * (0.5 * dxi * (fE[q] - shr1(f[q])));          // mlwhalf.f:33-35
        const double dfdx = dxi * (ff[q] - shr1(ff[q]));                   // mlwfull.f:30-33
        du = du - 0.5 * dt * dfdx;
        const double dEc = uE[q] - u[q];                                  // mlwartv.f:31-34
        const double dcW = shr1(dEc);
        du = du + dt * (dxi * (qxc * dxim * dEc - qxw * dxim1 * dcW));
        nu[q] = u[q] + du;
    }
    if (lane >= 1 && lane <= 62 && c <= ix - 1) {                         // mlw_m3.f:229-249
        const double nro = nu[0];
        const double nvx = nu[2] / nro;
        const double nvy = nu[3] / nro;
        const double nvz = nu[4] / nro;
        const double vv = nvx * nvx + nvy * nvy + nvz * nvz;
        const double bb = nu[5] * nu[5] + nu[6] * nu[6];
        a.out[0][p] = nro;
        a.out[1][p] = (gm - 1.0) * (nu[1] - 0.5 * nro * vv - pi8i * bb);
        a.out[2][p] = nvx;
        a.out[3][p] = nvy;
        a.out[4][p] = nvz;
        a.out[5][p] = nu[5];
        a.out[6][p] = nu[6];
    }
}

// ---------------------------------------------------------------- 2D
// fields: ro pr vx vy vz bx by bz az
constexpr int TX = 31, TY = 7;               // interior cells of a tile
constexpr int CX = TX + 2, CY = TY + 2;      // cells with the halo (33 x 9)
constexpr int NC = CX * CY;                  // 297
constexpr int KX = TX + 1;                   // corners per tile row (32 x 8 = 256)
// LDS rows of the flux block: fx of u0..u4 and u7, fy of the same, ez.
// fx(bx) = 0, fy(by) = 0, fy(bx) = ez, fx(by) = -ez (mlw_m3.f:178-201).
enum { FX0 = 0, FX7 = 5, FY0 = 6, FY7 = 11, FEZ = 12, NF = 13 };

// mlw_m3.f:107-202 for one cell (cell values and the half-step fluxes)
CANS_DEV void cell_values(double ro, double pr, double vx, double vy, double vz, double bx, double by, double bz,
                          double gm, double *u, double *fxo, double *fyo, double &ezo) {
    const double vv = vx * vx + vy * vy + vz * vz;
    const double bb = bx * bx + by * by + bz * bz;
    u[0] = ro;
    u[1] = pr / (gm - 1.0) + 0.5 * ro * vv + pi8i * bb;
    u[2] = ro * vx;
    u[3] = ro * vy;
    u[4] = ro * vz;
    u[5] = bx;
    u[6] = by;
    u[7] = bz;
    const double ex = -vy * bz + vz * by;
    const double ey = -vz * bx + vx * bz;
    const double ez = -vx * by + vy * bx;
    const double ep = pr * gm / (gm - 1.0) + 0.5 * ro * vv;
    fxo[0] = ro * vx;
    fyo[0] = ro * vy;
    fxo[1] = ep * vx + (bz * ey - by * ez) * pi4i;
    fyo[1] = ep * vy + (bx * ez - bz * ex) * pi4i;
    fxo[2] = ro * (vx * vx) + pr + pi8i * (by * by + bz * bz - bx * bx);
    fyo[2] = ro * vx * vy - pi4i * bx * by;
    fxo[3] = ro * vy * vx - pi4i * by * bx;
    fyo[3] = ro * (vy * vy) + pr + pi8i * (bx * bx + bz * bz - by * by);
    fxo[4] = ro * vz * vx - pi4i * bz * bx;
    fyo[4] = ro * vz * vy - pi4i * bz * by;
    fxo[5] = ey;      // slot 5 carries the bz fluxes (u index 7)
    fyo[5] = -ex;
    ezo = ez;
}

CANS_DEV double m3_update(const Mlw3Args &a, double dt, double *du, double daz, int lc, int li, int lj, int ci,
                          int cj, double (*sU)[NC], double (*sF)[NC], double (*sV)[NC], bool do_cfl);

__global__ __launch_bounds__(256) void mlw2d_m3_kernel(Mlw3Args a) {
    double dt;
    if (!m3_step(a, dt)) return;   // uniform across the launch
    __shared__ double red[4];
    __shared__ double sU[8][NC];
    __shared__ double sF[NF][NC];
    __shared__ double sV[2][NC];
    const int ix = a.ix, jx = a.jx;
    const int ntx = (ix - 2 + TX - 1) / TX;
    const int nty = (jx - 2 + TY - 1) / TY;
    const int b = xcd_remap((int)blockIdx.x, ntx * nty);
    const int i0 = 2 + TX * (b % ntx), j0 = 2 + TY * (b / ntx);   // first interior cell of the tile
    const int t = threadIdx.x;
    const double gm = a.gm;
    // (1) cells of the tile and its halo
    for (int lc = t; lc < NC; lc += 256) {
        const int li = lc % CX, lj = lc / CX;
        int i = i0 - 1 + li, j = j0 - 1 + lj;
        i = i > ix ? ix : i;
        j = j > jx ? jx : j;
        const size_t p = (size_t)(i - 1) + (size_t)ix * (size_t)(j - 1);
        double u[8], fx[6], fy[6], ez;
        const double vx = a.in[2][p], vy = a.in[3][p];
        cell_values(a.in[0][p], a.in[1][p], vx, vy, a.in[4][p], a.in[5][p], a.in[6][p], a.in[7][p], gm, u, fx, fy,
                    ez);
#pragma unroll
        for (int q = 0; q < 8; q++) sU[q][lc] = u[q];
#pragma unroll
        for (int q = 0; q < 6; q++) {
            sF[FX0 + q][lc] = fx[q];
            sF[FY0 + q][lc] = fy[q];
        }
        sF[FEZ][lc] = ez;
        sV[0][lc] = vx;
        sV[1][lc] = vy;
    }
    __syncthreads();
    // (2a) centred half-step term of the interior cell (mlwhalf.f, du part)
    const bool own = t < TX * TY;
    const int li = 1 + t % TX, lj = 1 + t / TX;
    const int ci = i0 - 1 + li, cj = j0 - 1 + lj;                 // this thread's cell (1-based)
    const int cic = ci > ix ? ix : ci, cjc = cj > jx ? jx : cj;
    const int lc = lj * CX + li;
    double du[8], daz = 0.0;
    if (own) {
        const double dxi = a.mx.di[cic - 1], dyi = a.my.di[cjc - 1];
        // slot of u index q in the flux block, and the ez-derived fluxes
#pragma unroll
        for (int q = 0; q < 8; q++) {
            double fxE, fxW, fyN, fyS;
            if (q < 5) {
                fxE = sF[FX0 + q][lc + 1]; fxW = sF[FX0 + q][lc - 1];
                fyN = sF[FY0 + q][lc + CX]; fyS = sF[FY0 + q][lc - CX];
            } else if (q == 5) {      // bx: fx = 0, fy = ez
                fxE = 0.0; fxW = 0.0;
                fyN = sF[FEZ][lc + CX]; fyS = sF[FEZ][lc - CX];
            } else if (q == 6) {      // by: fx = -ez, fy = 0
                fxE = -sF[FEZ][lc + 1]; fxW = -sF[FEZ][lc - 1];
                fyN = 0.0; fyS = 0.0;
            } else {                  // bz: fx = ey, fy = -ex
                fxE = sF[FX7][lc + 1]; fxW = sF[FX7][lc - 1];
                fyN = sF[FY7][lc + CX]; fyS = sF[FY7][lc - CX];
            }
            du[q] = 0.0;
            du[q] = du[q] - 0.5 * dt * (0.5 * dxi * (fxE - fxW)) - 0.5 * dt * (0.5 * dyi * (fyN - fyS));
        }
        daz = daz + 0.5 * dt * (-sF[FEZ][lc]);                      // mlwsrch.f, du part, s = -ez
    }
    // (2b) corner (i, j) between cells i, i+1 and j, j+1 (mlwhalf.f, un part)
    double cr[NF];
    {
        const int ki = t % KX, kj = t / KX;                         // 0..31, 0..7
        int i = i0 - 1 + ki, j = j0 - 1 + kj;
        i = i > ix ? ix : i;
        j = j > jx ? jx : j;
        const double dxim = a.mx.dim[i - 1], dyim = a.my.dim[j - 1];
        const int c00 = kj * CX + ki, c10 = c00 + 1, c01 = c00 + CX, c11 = c01 + 1;
        double un[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            double x00, x10, x01, x11, y00, y10, y01, y11;
            if (q < 5) {
                x00 = sF[FX0 + q][c00]; x10 = sF[FX0 + q][c10]; x01 = sF[FX0 + q][c01]; x11 = sF[FX0 + q][c11];
                y00 = sF[FY0 + q][c00]; y10 = sF[FY0 + q][c10]; y01 = sF[FY0 + q][c01]; y11 = sF[FY0 + q][c11];
            } else if (q == 5) {
                x00 = 0.0; x10 = 0.0; x01 = 0.0; x11 = 0.0;
                y00 = sF[FEZ][c00]; y10 = sF[FEZ][c10]; y01 = sF[FEZ][c01]; y11 = sF[FEZ][c11];
            } else if (q == 6) {
                x00 = -sF[FEZ][c00]; x10 = -sF[FEZ][c10]; x01 = -sF[FEZ][c01]; x11 = -sF[FEZ][c11];
                y00 = 0.0; y10 = 0.0; y01 = 0.0; y11 = 0.0;
            } else {
                x00 = sF[FX7][c00]; x10 = sF[FX7][c10]; x01 = sF[FX7][c01]; x11 = sF[FX7][c11];
                y00 = sF[FY7][c00]; y10 = sF[FY7][c10]; y01 = sF[FY7][c01]; y11 = sF[FY7][c11];
            }
            const double uh = 0.25 * (sU[q][c10] + sU[q][c00] + sU[q][c11] + sU[q][c01]);
            const double dfdx = dxim * (x10 - x00 + x11 - x01) / 2.0;
            const double dfdy = dyim * (y01 - y00 + y11 - y10) / 2.0;
            un[q] = uh - dt * dfdx - dt * dfdy;
        }
        // half-step primitives, E and corner fluxes (mlw_m3.f:214-310)
        const double roh = un[0], eeh = un[1], bxh = un[5], byh = un[6], bzh = un[7];
        const double vxh = un[2] / roh;
        const double vyh = un[3] / roh;
        const double vzh = un[4] / roh;
        const double vv = vxh * vxh + vyh * vyh + vzh * vzh;
        const double bb = bxh * bxh + byh * byh + bzh * bzh;
        const double prh = (gm - 1.0) * (eeh - 0.5 * roh * vv - pi8i * bb);
        const double exh = -vyh * bzh + vzh * byh;
        const double eyh = -vzh * bxh + vxh * bzh;
        const double ezh = -vxh * byh + vyh * bxh;
        const double ep = prh * gm / (gm - 1.0) + 0.5 * roh * vv;
        cr[FX0 + 0] = roh * vxh;
        cr[FY0 + 0] = roh * vyh;
        cr[FX0 + 1] = ep * vxh + (bzh * eyh - byh * ezh) * pi4i;
        cr[FY0 + 1] = ep * vyh + (bxh * ezh - bzh * exh) * pi4i;
        cr[FX0 + 2] = roh * (vxh * vxh) + prh + pi8i * (byh * byh + bzh * bzh - bxh * bxh);
        cr[FY0 + 2] = roh * vxh * vyh - pi4i * bxh * byh;
        cr[FX0 + 3] = roh * vyh * vxh - pi4i * byh * bxh;
        cr[FY0 + 3] = roh * (vyh * vyh) + prh + pi8i * (bxh * bxh + bzh * bzh - byh * byh);
        cr[FX0 + 4] = roh * vzh * vxh - pi4i * bzh * bxh;
        cr[FY0 + 4] = roh * vzh * vyh - pi4i * bzh * byh;
        cr[FX7] = eyh;
        cr[FY7] = -exh;
        cr[FEZ] = ezh;
    }
    __syncthreads();            // every reader of the cell fluxes is past this point
#pragma unroll
    for (int r = 0; r < NF; r++) sF[r][t] = cr[r];
    __syncthreads();
    // (3) full step, Az source, viscosity, update (mlw_m3.f:233-374); in a
    // periodic box the next step's cfl_m3 (cans2d/common/cfl_m3.f:44-62) of
    // the interior cells, every thread reaching the block minimum
    const int mg = a.per_margin;
    const bool do_cfl = a.clk && mg > 0 && a.safety > 0.0;
    double cq = kR4_1e20;
    if (own && ci <= ix - 1 && cj <= jx - 1) cq = m3_update(a, dt, du, daz, lc, li, lj, ci, cj, sU, sF, sV, do_cfl);
    if (do_cfl) block_min_to_global(cq, red, &a.clk->dtkey[(a.step + 1) % 3]);
}

// step (3) of mlw2d_m3_kernel for cell (ci, cj); with do_cfl, returns its
// cfl_m3 term when the cell is interior (else 1.e20)
CANS_DEV double m3_update(const Mlw3Args &a, double dt, double *du, double daz, int lc, int li, int lj, int ci,
                          int cj, double (*sU)[NC], double (*sF)[NC], double (*sV)[NC], bool do_cfl) {
    const int ix = a.ix, jx = a.jx;
    const double gm = a.gm;
    const int kSW = (lj - 1) * KX + (li - 1), kSE = kSW + 1, kNW = kSW + KX, kNE = kNW + 1;
    const double dxi = a.mx.di[ci - 1], dyi = a.my.di[cj - 1];
    const double ux0 = a.mx.u0[ci - 1], ux1 = a.mx.u1[ci - 1], uy0 = a.my.u0[cj - 1], uy1 = a.my.u1[cj - 1];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        double xSW, xSE, xNW, xNE, ySW, ySE, yNW, yNE;
        if (q < 5) {
            xSW = sF[FX0 + q][kSW]; xSE = sF[FX0 + q][kSE]; xNW = sF[FX0 + q][kNW]; xNE = sF[FX0 + q][kNE];
            ySW = sF[FY0 + q][kSW]; ySE = sF[FY0 + q][kSE]; yNW = sF[FY0 + q][kNW]; yNE = sF[FY0 + q][kNE];
        } else if (q == 5) {
            xSW = 0.0; xSE = 0.0; xNW = 0.0; xNE = 0.0;
            ySW = sF[FEZ][kSW]; ySE = sF[FEZ][kSE]; yNW = sF[FEZ][kNW]; yNE = sF[FEZ][kNE];
        } else if (q == 6) {
            xSW = -sF[FEZ][kSW]; xSE = -sF[FEZ][kSE]; xNW = -sF[FEZ][kNW]; xNE = -sF[FEZ][kNE];
            ySW = 0.0; ySE = 0.0; yNW = 0.0; yNE = 0.0;
        } else {
            xSW = sF[FX7][kSW]; xSE = sF[FX7][kSE]; xNW = sF[FX7][kNW]; xNE = sF[FX7][kNE];
            ySW = sF[FY7][kSW]; ySE = sF[FY7][kSE]; yNW = sF[FY7][kNW]; yNE = sF[FY7][kNE];
        }
        const double dfdx = dxi * (uy0 * (xSE - xSW) + uy1 * (xNE - xNW));   // mlwfull.f:31-41
        const double dfdy = dyi * (ux0 * (yNW - ySW) + ux1 * (yNE - ySE));
        du[q] = du[q] - 0.5 * dt * dfdx - 0.5 * dt * dfdy;
    }
    {   // mlwsrcf.f:33-40 with s = -ezh at the corners
        const double sh = uy0 * (ux0 * (-sF[FEZ][kSW]) + ux1 * (-sF[FEZ][kSE]))
                        + uy1 * (ux0 * (-sF[FEZ][kNW]) + ux1 * (-sF[FEZ][kNE]));
        daz = daz + 0.5 * dt * sh;
    }
    const double zero = 0.0;
    const double vx = sV[0][lc], vy = sV[1][lc];
    const double qxE = a.qav * a.dxm[ci - 1] * fmaxF(zero, __builtin_fabs(sV[0][lc + 1] - vx) - kR4_1em4);
    const double qxW = a.qav * a.dxm[ci - 2] * fmaxF(zero, __builtin_fabs(vx - sV[0][lc - 1]) - kR4_1em4);
    const double qyN = a.qav * a.dym[cj - 1] * fmaxF(zero, __builtin_fabs(sV[1][lc + CX] - vy) - kR4_1em4);
    const double qyS = a.qav * a.dym[cj - 2] * fmaxF(zero, __builtin_fabs(vy - sV[1][lc - CX]) - kR4_1em4);
    const double dxim = a.mx.dim[ci - 1], dxim1 = a.mx.dim[ci - 2];
    const double dyim = a.my.dim[cj - 1], dyim1 = a.my.dim[cj - 2];
    double nu[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {                                          // mlwartv.f:33-41
        const double uc = sU[q][lc];
        du[q] = du[q]
              + dt * (dxi * (qxE * dxim * (sU[q][lc + 1] - uc) - qxW * dxim1 * (uc - sU[q][lc - 1])))
              + dt * (dyi * (qyN * dyim * (sU[q][lc + CX] - uc) - qyS * dyim1 * (uc - sU[q][lc - CX])));
        nu[q] = uc + du[q];
    }
    const size_t p = (size_t)(ci - 1) + (size_t)ix * (size_t)(cj - 1);
    const double nro = nu[0];
    const double nvx = nu[2] / nro;
    const double nvy = nu[3] / nro;
    const double nvz = nu[4] / nro;
    const double vv = nvx * nvx + nvy * nvy + nvz * nvz;
    const double bb = nu[5] * nu[5] + nu[6] * nu[6] + nu[7] * nu[7];
    const double npr = (gm - 1.0) * (nu[1] - 0.5 * nro * vv - pi8i * bb);
    a.out[0][p] = nro;
    a.out[1][p] = npr;
    a.out[2][p] = nvx;
    a.out[3][p] = nvy;
    a.out[4][p] = nvz;
    a.out[5][p] = nu[5];
    a.out[6][p] = nu[6];
    a.out[7][p] = nu[7];
    a.out[8][p] = a.in[8][p] + daz;
    const int mg = a.per_margin;
    if (do_cfl && ci >= mg + 1 && ci <= ix - mg && cj >= mg + 1 && cj <= jx - mg) {
        // v2 and b2 of cfl_m3 are vv and bb (the same sums, z components last)
        const double dg = fminF(image_min(a.dx, ci, ix, mg), image_min(a.dy, cj, jx, mg));
        return cfl2_term(a.safety, gm, nro, npr, vv, bb, dg);
    }
    return kR4_1e20;
}

}  // namespace

hipError_t launch_mlw_m3(const Mlw3Args &a, int dim, hipStream_t s) {
    if (dim == 1) {
        const int nseg = (a.ix - 2 + 61) / 62;
        hipLaunchKernelGGL(mlw1d_m3_kernel, dim3((nseg + 3) / 4), dim3(256), 0, s, a);
    } else {
        const int ntx = (a.ix - 2 + TX - 1) / TX, nty = (a.jx - 2 + TY - 1) / TY;
        hipLaunchKernelGGL(mlw2d_m3_kernel, dim3(ntx * nty), dim3(256), 0, s, a);
    }
    return hipGetLastError();
}


// ---------------------------------------------------------------- 2D cylindrical
// mlw_m_c (cans2d/hdmlw/mlw_m_c.f:2-350): fields ro pr vx vz bx bz ay on the
// (r, z) = (x, z) plane.  The same one-pass LDS tile as mlw2d_m3_kernel; in
// addition every variable but bx carries the geometric source -F_x/r, which
// enters the corner half step (mlwsrch.f) and the cell (mlwsrch.f, cell
// value; mlwsrcf.f, corner values).
namespace {

// LDS rows: cell u (6), x and z fluxes of ro ee rx rz, Ey, cell sources of
// ro ee rx rz bz, vx, vz.  Corner values overwrite rows CF0..CSBZ.
enum { CU0 = 0, CFX = 6, CFZ = 10, CEY = 14, CS = 15, CSBZ = 19, CVX = 20, CVZ = 21, CNR = 22 };

CANS_DEV void cyl_fluxes(double ro, double pr, double vx, double vz, double bx, double bz, double ey, double r,
                         double gm, double *fx, double *fz, double *ss) {
    fx[0] = ro * vx;
    fz[0] = ro * vz;
    ss[0] = -fx[0] / r;
    const double vv = vx * vx + vz * vz;
    const double ep = pr * gm / (gm - 1.0) + 0.5 * ro * vv;
    fx[1] = ep * vx + (bz * ey) * pi4i;
    fz[1] = ep * vz + (-bx * ey) * pi4i;
    ss[1] = -fx[1] / r;
    fx[2] = ro * (vx * vx) + pr + pi8i * (bz * bz - bx * bx);
    fz[2] = ro * vx * vz - pi4i * bx * bz;
    ss[2] = -(ro * (vx * vx) + pi4i * (-(bx * bx))) / r;
    fx[3] = ro * vz * vx - pi4i * bz * bx;
    fz[3] = ro * (vz * vz) + pr + pi8i * (bx * bx - bz * bz);
    ss[3] = -fx[3] / r;
}

__global__ __launch_bounds__(256) void mlw2d_c_kernel(Mlw3Args a) {
    double dt;
    if (!m3_step(a, dt)) return;
    __shared__ double sA[CNR][NC];
    const int ix = a.ix, jx = a.jx;
    const int ntx = (ix - 2 + TX - 1) / TX;
    const int nty = (jx - 2 + TY - 1) / TY;
    const int b = xcd_remap((int)blockIdx.x, ntx * nty);
    const int i0 = 2 + TX * (b % ntx), j0 = 2 + TY * (b / ntx);
    const int t = threadIdx.x;
    const double gm = a.gm;
    // (1) cells of the tile and its halo (mlw_m_c.f:105-198)
    for (int lc = t; lc < NC; lc += 256) {
        const int li = lc % CX, lj = lc / CX;
        int i = i0 - 1 + li, j = j0 - 1 + lj;
        i = i > ix ? ix : i;
        j = j > jx ? jx : j;
        const size_t p = (size_t)(i - 1) + (size_t)ix * (size_t)(j - 1);
        const double ro = a.in[0][p], pr = a.in[1][p], vx = a.in[2][p], vz = a.in[3][p];
        const double bx = a.in[4][p], bz = a.in[5][p];
        const double vv = vx * vx + vz * vz;
        const double bb = bx * bx + bz * bz;
        sA[CU0 + 0][lc] = ro;
        sA[CU0 + 1][lc] = pr / (gm - 1.0) + 0.5 * ro * vv + pi8i * bb;
        sA[CU0 + 2][lc] = ro * vx;
        sA[CU0 + 3][lc] = ro * vz;
        sA[CU0 + 4][lc] = bx;
        sA[CU0 + 5][lc] = bz;
        const double ey = -vz * bx + vx * bz;
        const double r = a.x[i - 1];
        double fx[4], fz[4], ss[4];
        cyl_fluxes(ro, pr, vx, vz, bx, bz, ey, r, gm, fx, fz, ss);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            sA[CFX + q][lc] = fx[q];
            sA[CFZ + q][lc] = fz[q];
            sA[CS + q][lc] = ss[q];
        }
        sA[CEY][lc] = ey;
        sA[CSBZ][lc] = -ey / r;          // bz: fx = ey, ss = -fx/x
        sA[CVX][lc] = vx;
        sA[CVZ][lc] = vz;
    }
    __syncthreads();
    // (2a) centred half-step term and cell sources (mlwhalf.f, mlwsrch.f)
    const bool own = t < TX * TY;
    const int li = 1 + t % TX, lj = 1 + t / TX;
    const int ci = i0 - 1 + li, cj = j0 - 1 + lj;
    const int cic = ci > ix ? ix : ci, cjc = cj > jx ? jx : cj;
    const int lc = lj * CX + li;
    double du[6], day = 0.0;
    if (own) {
        const double dxi = a.mx.di[cic - 1], dzi = a.my.di[cjc - 1];
#pragma unroll
        for (int q = 0; q < 6; q++) {
            double fxE, fxW, fzN, fzS;
            if (q < 4) {
                fxE = sA[CFX + q][lc + 1]; fxW = sA[CFX + q][lc - 1];
                fzN = sA[CFZ + q][lc + CX]; fzS = sA[CFZ + q][lc - CX];
            } else if (q == 4) {      // bx: fx = 0, fz = -ey
                fxE = 0.0; fxW = 0.0;
                fzN = -sA[CEY][lc + CX]; fzS = -sA[CEY][lc - CX];
            } else {                  // bz: fx = ey, fz = 0
                fxE = sA[CEY][lc + 1]; fxW = sA[CEY][lc - 1];
                fzN = 0.0; fzS = 0.0;
            }
            du[q] = 0.0;
            du[q] = du[q] - 0.5 * dt * (0.5 * dxi * (fxE - fxW)) - 0.5 * dt * (0.5 * dzi * (fzN - fzS));
            if (q < 4) du[q] = du[q] + 0.5 * dt * sA[CS + q][lc];
            if (q == 5) du[q] = du[q] + 0.5 * dt * sA[CSBZ][lc];
        }
        day = day + 0.5 * dt * (-sA[CEY][lc]);
    }
    // (2b) corner (i, j): half step with the sources, primitives, fluxes
    double cr[CSBZ - CFX + 1];
    {
        const int ki = t % KX, kj = t / KX;
        int i = i0 - 1 + ki, j = j0 - 1 + kj;
        i = i > ix ? ix : i;
        j = j > jx ? jx : j;
        const double dxim = a.mx.dim[i - 1], dzim = a.my.dim[j - 1];
        const int c00 = kj * CX + ki, c10 = c00 + 1, c01 = c00 + CX, c11 = c01 + 1;
        double un[6];
#pragma unroll
        for (int q = 0; q < 6; q++) {
            double x00, x10, x01, x11, z00, z10, z01, z11;
            if (q < 4) {
                x00 = sA[CFX + q][c00]; x10 = sA[CFX + q][c10]; x01 = sA[CFX + q][c01]; x11 = sA[CFX + q][c11];
                z00 = sA[CFZ + q][c00]; z10 = sA[CFZ + q][c10]; z01 = sA[CFZ + q][c01]; z11 = sA[CFZ + q][c11];
            } else if (q == 4) {
                x00 = 0.0; x10 = 0.0; x01 = 0.0; x11 = 0.0;
                z00 = -sA[CEY][c00]; z10 = -sA[CEY][c10]; z01 = -sA[CEY][c01]; z11 = -sA[CEY][c11];
            } else {
                x00 = sA[CEY][c00]; x10 = sA[CEY][c10]; x01 = sA[CEY][c01]; x11 = sA[CEY][c11];
                z00 = 0.0; z10 = 0.0; z01 = 0.0; z11 = 0.0;
            }
            const double uh = 0.25 * (sA[CU0 + q][c10] + sA[CU0 + q][c00] + sA[CU0 + q][c11] + sA[CU0 + q][c01]);
            const double dfdx = dxim * (x10 - x00 + x11 - x01) / 2.0;
            const double dfdz = dzim * (z01 - z00 + z11 - z10) / 2.0;
            un[q] = uh - dt * dfdx - dt * dfdz;
            if (q != 4) {   // mlwsrch.f:38-44, every variable but bx
                const int rs = q < 4 ? CS + q : CSBZ;
                const double sh = (sA[rs][c10] + sA[rs][c00] + sA[rs][c11] + sA[rs][c01]) / 4.0;
                un[q] = un[q] + dt * sh;
            }
        }
        const double roh = un[0], eeh = un[1], bxh = un[4], bzh = un[5];
        const double vxh = un[2] / roh;
        const double vzh = un[3] / roh;
        const double vv = vxh * vxh + vzh * vzh;
        const double bb = bxh * bxh + bzh * bzh;
        const double prh = (gm - 1.0) * (eeh - 0.5 * roh * vv - pi8i * bb);
        const double eyh = -vzh * bxh + vxh * bzh;
        double fx[4], fz[4], ss[4];
        cyl_fluxes(roh, prh, vxh, vzh, bxh, bzh, eyh, a.xm[i - 1], gm, fx, fz, ss);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            cr[CFX - CFX + q] = fx[q];
            cr[CFZ - CFX + q] = fz[q];
            cr[CS - CFX + q] = ss[q];
        }
        cr[CEY - CFX] = eyh;
        cr[CSBZ - CFX] = -eyh / a.x[i - 1];   // mlw_m_c.f:284 divides the bz source by x, not xm
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r <= CSBZ - CFX; r++) sA[CFX + r][t] = cr[r];
    __syncthreads();
    // (3) full step, corner sources, viscosity, update (mlw_m_c.f:219-346)
    if (!own || ci > ix - 1 || cj > jx - 1) return;
    const int kSW = (lj - 1) * KX + (li - 1), kSE = kSW + 1, kNW = kSW + KX, kNE = kNW + 1;
    const double dxi = a.mx.di[ci - 1], dzi = a.my.di[cj - 1];
    const double ux0 = a.mx.u0[ci - 1], ux1 = a.mx.u1[ci - 1], uz0 = a.my.u0[cj - 1], uz1 = a.my.u1[cj - 1];
#pragma unroll
    for (int q = 0; q < 6; q++) {
        double xSW, xSE, xNW, xNE, zSW, zSE, zNW, zNE;
        if (q < 4) {
            xSW = sA[CFX + q][kSW]; xSE = sA[CFX + q][kSE]; xNW = sA[CFX + q][kNW]; xNE = sA[CFX + q][kNE];
            zSW = sA[CFZ + q][kSW]; zSE = sA[CFZ + q][kSE]; zNW = sA[CFZ + q][kNW]; zNE = sA[CFZ + q][kNE];
        } else if (q == 4) {
            xSW = 0.0; xSE = 0.0; xNW = 0.0; xNE = 0.0;
            zSW = -sA[CEY][kSW]; zSE = -sA[CEY][kSE]; zNW = -sA[CEY][kNW]; zNE = -sA[CEY][kNE];
        } else {
            xSW = sA[CEY][kSW]; xSE = sA[CEY][kSE]; xNW = sA[CEY][kNW]; xNE = sA[CEY][kNE];
            zSW = 0.0; zSE = 0.0; zNW = 0.0; zNE = 0.0;
        }
        const double dfdx = dxi * (uz0 * (xSE - xSW) + uz1 * (xNE - xNW));
        const double dfdz = dzi * (ux0 * (zNW - zSW) + ux1 * (zNE - zSE));
        du[q] = du[q] - 0.5 * dt * dfdx - 0.5 * dt * dfdz;
        if (q != 4) {   // mlwsrcf.f:33-40
            const int rs = q < 4 ? CS + q : CSBZ;
            const double sh = uz0 * (ux0 * sA[rs][kSW] + ux1 * sA[rs][kSE]) + uz1 * (ux0 * sA[rs][kNW] + ux1 * sA[rs][kNE]);
            du[q] = du[q] + 0.5 * dt * sh;
        }
    }
    {
        const double sh = uz0 * (ux0 * (-sA[CEY][kSW]) + ux1 * (-sA[CEY][kSE]))
                        + uz1 * (ux0 * (-sA[CEY][kNW]) + ux1 * (-sA[CEY][kNE]));
        day = day + 0.5 * dt * sh;
    }
    const double zero = 0.0;
    const double vx = sA[CVX][lc], vz = sA[CVZ][lc];
    const double qxE = a.qav * a.dxm[ci - 1] * fmaxF(zero, __builtin_fabs(sA[CVX][lc + 1] - vx) - kR4_1em4);
    const double qxW = a.qav * a.dxm[ci - 2] * fmaxF(zero, __builtin_fabs(vx - sA[CVX][lc - 1]) - kR4_1em4);
    const double qzN = a.qav * a.dym[cj - 1] * fmaxF(zero, __builtin_fabs(sA[CVZ][lc + CX] - vz) - kR4_1em4);
    const double qzS = a.qav * a.dym[cj - 2] * fmaxF(zero, __builtin_fabs(vz - sA[CVZ][lc - CX]) - kR4_1em4);
    const double dxim = a.mx.dim[ci - 1], dxim1 = a.mx.dim[ci - 2];
    const double dzim = a.my.dim[cj - 1], dzim1 = a.my.dim[cj - 2];
    double nu[6];
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const double uc = sA[CU0 + q][lc];
        du[q] = du[q]
              + dt * (dxi * (qxE * dxim * (sA[CU0 + q][lc + 1] - uc) - qxW * dxim1 * (uc - sA[CU0 + q][lc - 1])))
              + dt * (dzi * (qzN * dzim * (sA[CU0 + q][lc + CX] - uc) - qzS * dzim1 * (uc - sA[CU0 + q][lc - CX])));
        nu[q] = uc + du[q];
    }
    const size_t p = (size_t)(ci - 1) + (size_t)ix * (size_t)(cj - 1);
    const double nro = nu[0];
    const double nvx = nu[2] / nro;
    const double nvz = nu[3] / nro;
    const double vv = nvx * nvx + nvz * nvz;
    const double bb = nu[4] * nu[4] + nu[5] * nu[5];
    a.out[0][p] = nro;
    a.out[1][p] = (gm - 1.0) * (nu[1] - 0.5 * nro * vv - pi8i * bb);
    a.out[2][p] = nvx;
    a.out[3][p] = nvz;
    a.out[4][p] = nu[4];
    a.out[5][p] = nu[5];
    a.out[6][p] = a.in[6][p] + day;
}

}  // namespace

hipError_t launch_mlw_m_c(const Mlw3Args &a, hipStream_t s) {
    const int ntx = (a.ix - 2 + TX - 1) / TX, nty = (a.jx - 2 + TY - 1) / TY;
    hipLaunchKernelGGL(mlw2d_c_kernel, dim3(ntx * nty), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace cans
