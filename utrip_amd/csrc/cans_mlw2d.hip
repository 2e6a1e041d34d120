// cans_mlw2d.hip -- fused 2D modified Lax-Wendroff MHD step (gfx950).
//
// One launch does the whole of cans2d/hdmlw/mlw_m.f:2-324 (mlwhalf, mlwsrch,
// mlwfull, mlwsrcf, mlwartv, update, primitive recovery) for a tile of
// TX x TY interior cells, and optionally the periodic ghost fill of
// cans2d/bc/bdper{x,y}.f and the CFL reduction of cans2d/common/cfl_m.f for
// the next step.  State is structure-of-arrays FP64, column-major (ix,jx)
// exactly as the reference holds it, so the kernel reads each field once
// with coalesced 8-byte lanes along x.
//
// Per block:
//   stage 1  load the (TX+2)x(TY+2) halo cells' primitives, form the
//            conserved variables and the cell-centred fluxes into LDS;
//   stage 2  the (TX+1)x(TY+1) corner half-step states (mlwhalf.f:45-52)
//            and their fluxes (mlw_m.f:189-262) into registers, while the
//            centre term of mlwhalf (mlwhalf.f:36-41) goes into du;
//   stage 3  corner fluxes to LDS (aliasing the cell fluxes), full step,
//            sources, artificial viscosity, update and primitives.
// Every value is formed by the reference's own expression in its own order,
// so a value computed by two blocks is the same bits in both.
#include "cans_common.hpp"
#include "cans_clock.hpp"
#include "cans_kernels.hpp"

namespace cans {
namespace {

constexpr int TX = 32, TY = 8;
constexpr int NT = TX * TY;                  // 256 threads, one interior cell each
constexpr int HX = TX + 2, HY = TY + 2;      // halo cells
constexpr int NH = HX * HY;                  // 340
constexpr int CXN = TX + 1, CYN = TY + 1;    // corners
constexpr int NC = CXN * CYN;                // 297

constexpr double pi8i = 1.0 / kPi / 8.0;     // mlw_m.f:27  pi8i=1./pi/8.
constexpr double pi4i = 1.0 / kPi / 4.0;     // mlw_m.f:28  pi4i=1./pi/4.

// cell arrays in LDS
enum { C_RO, C_EE, C_RX, C_RY, C_BX, C_BY, C_VX, C_VY, NCELL };
// cell fluxes (aliased by the corner fluxes after stage 2)
enum { F_EEX, F_EEY, F_RXX, F_RXY, F_RYX, F_RYY, F_EZ, NFLUX };
// corner fluxes
enum { H_ROX, H_ROY, H_EEX, H_EEY, H_RXX, H_RXY, H_RYX, H_RYY, H_EZ, NCORN };

constexpr int kFluxRegion = (NFLUX * NH > NCORN * NC) ? NFLUX * NH : NCORN * NC;

struct CornerOut {
    double v[NCORN];
};

__global__ __launch_bounds__(NT) void mlw2d_kernel(Mlw2dArgs a) {
    __shared__ double cell[NCELL * NH];
    __shared__ double flx[kFluxRegion];
    __shared__ double red[NT / 64];

    const int ix = a.ix, jx = a.jx;
    const int ntx = (ix - 2 + TX - 1) / TX;
    const int nty = (jx - 2 + TY - 1) / TY;
    const int tile = xcd_remap(blockIdx.x, ntx * nty);
    const int i0 = 2 + (tile % ntx) * TX;    // first interior cell of the tile (1-based)
    const int j0 = 2 + (tile / ntx) * TY;
    const int tid = threadIdx.x;

    double dt = a.dt;
    bool run = true;
    if (a.clk) {
        StepCtl c = clock_use(a.clk, a.step, a.peek != 0, blockIdx.x == 0 && tid == 0);
        run = c.run;
        dt = c.dt;
    }
    if (!run) return;
    const double gm = a.gm;

    // ---- stage 1: halo cells -> conserved variables and fluxes (mlw_m.f:101-176)
    for (int c = tid; c < NH; c += NT) {
        const int hx = c % HX, hy = c / HX;
        const int i = i0 - 1 + hx, j = j0 - 1 + hy;
        if (i <= ix && j <= jx) {
            const size_t p = (size_t)(i - 1) + (size_t)ix * (size_t)(j - 1);
            const double ro = a.ro[p], pr = a.pr[p], vx = a.vx[p], vy = a.vy[p];
            const double bx = a.bx[p], by = a.by[p];
            const double vv = vx * vx + vy * vy;
            const double bb = bx * bx + by * by;
            const double ee = pr / (gm - 1.0) + 0.5 * ro * vv + pi8i * bb;
            const double ez = -vx * by + vy * bx;
            const double ep = pr * gm / (gm - 1.0) + 0.5 * ro * vv;
            cell[C_RO * NH + c] = ro;
            cell[C_EE * NH + c] = ee;
            cell[C_RX * NH + c] = ro * vx;
            cell[C_RY * NH + c] = ro * vy;
            cell[C_BX * NH + c] = bx;
            cell[C_BY * NH + c] = by;
            cell[C_VX * NH + c] = vx;
            cell[C_VY * NH + c] = vy;
            flx[F_EEX * NH + c] = ep * vx - by * ez * pi4i;
            flx[F_EEY * NH + c] = ep * vy + bx * ez * pi4i;
            flx[F_RXX * NH + c] = ro * (vx * vx) + pr + pi8i * (by * by - bx * bx);
            flx[F_RXY * NH + c] = ro * vx * vy - pi4i * bx * by;
            flx[F_RYX * NH + c] = ro * vy * vx - pi4i * by * bx;
            flx[F_RYY * NH + c] = ro * (vy * vy) + pr + pi8i * (bx * bx - by * by);
            flx[F_EZ * NH + c] = ez;
        }
    }
    __syncthreads();

    // ---- stage 2a: centre term of mlwhalf for the own cell (mlwhalf.f:36-41)
    const int tx = tid % TX, ty = tid / TX;
    const int i = i0 + tx, j = j0 + ty;
    const bool own = (i <= ix - 1) && (j <= jx - 1);
    const int hc = (tx + 1) + HX * (ty + 1);          // own cell in the halo
    const int hE = hc + 1, hW = hc - 1, hN = hc + HX, hS = hc - HX;
    double du[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};    // ro ee rx ry bx by
    double daz = 0.0;
    double dxi = 0, dxim = 0, dxim1 = 0, dyi = 0, dyim = 0, dyim1 = 0, ux0 = 0, ux1 = 0, uy0 = 0, uy1 = 0;
    if (own) {
        dxi = a.mx.di[i - 1]; dxim = a.mx.dim[i - 1]; dxim1 = a.mx.dim[i - 2];
        ux0 = a.mx.u0[i - 1]; ux1 = a.mx.u1[i - 1];
        dyi = a.my.di[j - 1]; dyim = a.my.dim[j - 1]; dyim1 = a.my.dim[j - 2];
        uy0 = a.my.u0[j - 1]; uy1 = a.my.u1[j - 1];
        // fluxes of each equation at E/W (x) and N/S (y)
        const double fxE[6] = {cell[C_RX * NH + hE], flx[F_EEX * NH + hE], flx[F_RXX * NH + hE],
                               flx[F_RYX * NH + hE], 0.0, -flx[F_EZ * NH + hE]};
        const double fxW[6] = {cell[C_RX * NH + hW], flx[F_EEX * NH + hW], flx[F_RXX * NH + hW],
                               flx[F_RYX * NH + hW], 0.0, -flx[F_EZ * NH + hW]};
        const double fyN[6] = {cell[C_RY * NH + hN], flx[F_EEY * NH + hN], flx[F_RXY * NH + hN],
                               flx[F_RYY * NH + hN], flx[F_EZ * NH + hN], 0.0};
        const double fyS[6] = {cell[C_RY * NH + hS], flx[F_EEY * NH + hS], flx[F_RXY * NH + hS],
                               flx[F_RYY * NH + hS], flx[F_EZ * NH + hS], 0.0};
#pragma unroll
        for (int q = 0; q < 6; q++)
            du[q] = du[q] - 0.5 * dt * (0.5 * dxi * (fxE[q] - fxW[q]))
                          - 0.5 * dt * (0.5 * dyi * (fyN[q] - fyS[q]));
        // mlwsrch.f:32-36, s = -ez
        daz = daz + 0.5 * dt * (-flx[F_EZ * NH + hc]);
    }

    // ---- stage 2b: corner half-step states and fluxes (mlwhalf.f:45-52, mlw_m.f:189-262)
    CornerOut co[(NC + NT - 1) / NT];
#pragma unroll
    for (int r = 0; r < (NC + NT - 1) / NT; r++) {
        const int c = tid + r * NT;
        if (c >= NC) break;
        const int cx = c % CXN, cy = c / CXN;
        const int ci = i0 - 1 + cx, cj = j0 - 1 + cy;   // corner between cells ci,ci+1 / cj,cj+1
        if (ci > ix - 1 || cj > jx - 1) continue;
        const int h00 = cx + HX * cy, h10 = h00 + 1, h01 = h00 + HX, h11 = h01 + 1;
        const double cdxim = a.mx.dim[ci - 1], cdyim = a.my.dim[cj - 1];
        double un[6];
#pragma unroll
        for (int q = 0; q < 6; q++) {
            const double *u = cell + q * NH;   // C_RO..C_BY are 0..5 in conserved order
            double fx00, fx10, fx01, fx11, fy00, fy10, fy01, fy11;
            switch (q) {
            case 0:
                fx00 = cell[C_RX * NH + h00]; fx10 = cell[C_RX * NH + h10]; fx01 = cell[C_RX * NH + h01]; fx11 = cell[C_RX * NH + h11];
                fy00 = cell[C_RY * NH + h00]; fy10 = cell[C_RY * NH + h10]; fy01 = cell[C_RY * NH + h01]; fy11 = cell[C_RY * NH + h11];
                break;
            case 1:
                fx00 = flx[F_EEX * NH + h00]; fx10 = flx[F_EEX * NH + h10]; fx01 = flx[F_EEX * NH + h01]; fx11 = flx[F_EEX * NH + h11];
                fy00 = flx[F_EEY * NH + h00]; fy10 = flx[F_EEY * NH + h10]; fy01 = flx[F_EEY * NH + h01]; fy11 = flx[F_EEY * NH + h11];
                break;
            case 2:
                fx00 = flx[F_RXX * NH + h00]; fx10 = flx[F_RXX * NH + h10]; fx01 = flx[F_RXX * NH + h01]; fx11 = flx[F_RXX * NH + h11];
                fy00 = flx[F_RXY * NH + h00]; fy10 = flx[F_RXY * NH + h10]; fy01 = flx[F_RXY * NH + h01]; fy11 = flx[F_RXY * NH + h11];
                break;
            case 3:
                fx00 = flx[F_RYX * NH + h00]; fx10 = flx[F_RYX * NH + h10]; fx01 = flx[F_RYX * NH + h01]; fx11 = flx[F_RYX * NH + h11];
                fy00 = flx[F_RYY * NH + h00]; fy10 = flx[F_RYY * NH + h10]; fy01 = flx[F_RYY * NH + h01]; fy11 = flx[F_RYY * NH + h11];
                break;
            case 4:  // bx: fx = 0, fy = ez
                fx00 = 0.0; fx10 = 0.0; fx01 = 0.0; fx11 = 0.0;
                fy00 = flx[F_EZ * NH + h00]; fy10 = flx[F_EZ * NH + h10]; fy01 = flx[F_EZ * NH + h01]; fy11 = flx[F_EZ * NH + h11];
                break;
            default: // by: fx = -ez, fy = 0
                fx00 = -flx[F_EZ * NH + h00]; fx10 = -flx[F_EZ * NH + h10]; fx01 = -flx[F_EZ * NH + h01]; fx11 = -flx[F_EZ * NH + h11];
                fy00 = 0.0; fy10 = 0.0; fy01 = 0.0; fy11 = 0.0;
                break;
            }
            const double uh = 0.25 * (u[h10] + u[h00] + u[h11] + u[h01]);
            const double dfdx = cdxim * (fx10 - fx00 + fx11 - fx01) / 2.0;
            const double dfdy = cdyim * (fy01 - fy00 + fy11 - fy10) / 2.0;
            un[q] = uh - dt * dfdx - dt * dfdy;
        }
        // half-step primitives and fluxes (mlw_m.f:189-262)
        const double roh = un[0], eeh = un[1], rxh = un[2], ryh = un[3], bxh = un[4], byh = un[5];
        const double vxh = rxh / roh;
        const double vyh = ryh / roh;
        const double vv = vxh * vxh + vyh * vyh;
        const double bb = bxh * bxh + byh * byh;
        const double prh = (gm - 1.0) * (eeh - 0.5 * roh * vv - pi8i * bb);
        const double ezh = -vxh * byh + vyh * bxh;
        const double ep = prh * gm / (gm - 1.0) + 0.5 * roh * vv;
        co[r].v[H_ROX] = roh * vxh;
        co[r].v[H_ROY] = roh * vyh;
        co[r].v[H_EEX] = ep * vxh - byh * ezh * pi4i;
        co[r].v[H_EEY] = ep * vyh + bxh * ezh * pi4i;
        co[r].v[H_RXX] = roh * (vxh * vxh) + prh + pi8i * (byh * byh - bxh * bxh);
        co[r].v[H_RXY] = roh * vxh * vyh - pi4i * bxh * byh;
        co[r].v[H_RYX] = roh * vyh * vxh - pi4i * byh * bxh;
        co[r].v[H_RYY] = roh * (vyh * vyh) + prh + pi8i * (bxh * bxh - byh * byh);
        co[r].v[H_EZ] = ezh;
    }
    __syncthreads();   // all reads of the cell fluxes are done
#pragma unroll
    for (int r = 0; r < (NC + NT - 1) / NT; r++) {
        const int c = tid + r * NT;
        if (c >= NC) break;
#pragma unroll
        for (int v = 0; v < NCORN; v++) flx[v * NC + c] = co[r].v[v];
    }
    __syncthreads();

    double dtmin_local = kR4_1e20;
    const bool do_cfl = a.clk && a.safety > 0.0;
    if (own) {
        // ---- stage 3: full step (mlwfull.f:31-41), mlwsrcf.f:33-40, AV, update
        const int cSW = tx + CXN * ty;        // corner (i-1, j-1)
        const int cSE = cSW + 1;              // corner (i,   j-1)
        const int cNW = cSW + CXN;            // corner (i-1, j)
        const int cNE = cNW + 1;              // corner (i,   j)
        const int fxid[6] = {H_ROX, H_EEX, H_RXX, H_RYX, -1, H_EZ};
        const int fyid[6] = {H_ROY, H_EEY, H_RXY, H_RYY, H_EZ, -1};
#pragma unroll
        for (int q = 0; q < 6; q++) {
            double fSW, fSE, fNW, fNE, gSW, gSE, gNW, gNE;
            if (fxid[q] < 0) {
                fSW = fSE = fNW = fNE = 0.0;
            } else {
                fSW = flx[fxid[q] * NC + cSW]; fSE = flx[fxid[q] * NC + cSE];
                fNW = flx[fxid[q] * NC + cNW]; fNE = flx[fxid[q] * NC + cNE];
                if (q == 5) { fSW = -fSW; fSE = -fSE; fNW = -fNW; fNE = -fNE; }  // by: fx = -ezh
            }
            if (fyid[q] < 0) {
                gSW = gSE = gNW = gNE = 0.0;
            } else {
                gSW = flx[fyid[q] * NC + cSW]; gSE = flx[fyid[q] * NC + cSE];
                gNW = flx[fyid[q] * NC + cNW]; gNE = flx[fyid[q] * NC + cNE];
            }
            const double dfdx = dxi * (uy0 * (fSE - fSW) + uy1 * (fNE - fNW));
            const double dfdy = dyi * (ux0 * (gNW - gSW) + ux1 * (gNE - gSE));
            du[q] = du[q] - 0.5 * dt * dfdx - 0.5 * dt * dfdy;
        }
        {
            const double sSW = -flx[H_EZ * NC + cSW], sSE = -flx[H_EZ * NC + cSE];
            const double sNW = -flx[H_EZ * NC + cNW], sNE = -flx[H_EZ * NC + cNE];
            const double sh = uy0 * (ux0 * sSW + ux1 * sSE) + uy1 * (ux0 * sNW + ux1 * sNE);
            daz = daz + 0.5 * dt * sh;
        }
        // artificial viscosity (mlw_m.f:276-295, mlwartv.f:33-41)
        {
            const double zero = 0.0;
            const double vxc = cell[C_VX * NH + hc], vyc = cell[C_VY * NH + hc];
            const double qxc = a.qav * a.dxm[i - 1] * fmaxF(zero, __builtin_fabs(cell[C_VX * NH + hE] - vxc) - kR4_1em4);
            const double qxw = a.qav * a.dxm[i - 2] * fmaxF(zero, __builtin_fabs(vxc - cell[C_VX * NH + hW]) - kR4_1em4);
            const double qyc = a.qav * a.dym[j - 1] * fmaxF(zero, __builtin_fabs(cell[C_VY * NH + hN] - vyc) - kR4_1em4);
            const double qys = a.qav * a.dym[j - 2] * fmaxF(zero, __builtin_fabs(vyc - cell[C_VY * NH + hS]) - kR4_1em4);
#pragma unroll
            for (int q = 0; q < 6; q++) {
                const double *u = cell + q * NH;
                const double uc = u[hc];
                du[q] = du[q]
                      + dt * (dxi * (qxc * dxim * (u[hE] - uc) - qxw * dxim1 * (uc - u[hW])))
                      + dt * (dyi * (qyc * dyim * (u[hN] - uc) - qys * dyim1 * (uc - u[hS])));
            }
        }
        // update and primitive recovery (mlw_m.f:299-321)
        const size_t p = (size_t)(i - 1) + (size_t)ix * (size_t)(j - 1);
        const double ro = cell[C_RO * NH + hc] + du[0];
        const double ee = cell[C_EE * NH + hc] + du[1];
        const double rx = cell[C_RX * NH + hc] + du[2];
        const double ry = cell[C_RY * NH + hc] + du[3];
        const double bx = cell[C_BX * NH + hc] + du[4];
        const double by = cell[C_BY * NH + hc] + du[5];
        const double az = a.az[p] + daz;
        const double vx = rx / ro;
        const double vy = ry / ro;
        const double vv = vx * vx + vy * vy;
        const double bb = bx * bx + by * by;
        const double pr = (gm - 1.0) * (ee - 0.5 * ro * vv - pi8i * bb);

        // the cell and, in a periodic box, its ghost images: after bdperx
        // then bdpery, ghost (gi,gj) holds interior (src_x(gi), src_y(gj))
        const int m = a.per_margin;   // the host checks ix-2m >= m and jx-2m >= m
        int xi[3] = {i, 0, 0}, yj[3] = {j, 0, 0};
        int nxi = 1, nyj = 1;
        if (m > 0) {
            if (i >= ix - 2 * m + 1) xi[nxi++] = i - (ix - 2 * m);   // low ghost  <- high interior
            if (i <= 2 * m) xi[nxi++] = i + (ix - 2 * m);            // high ghost <- low interior
            if (j >= jx - 2 * m + 1) yj[nyj++] = j - (jx - 2 * m);
            if (j <= 2 * m) yj[nyj++] = j + (jx - 2 * m);
        }
        for (int b = 0; b < nyj; b++) {
            for (int c = 0; c < nxi; c++) {
                const size_t q = (size_t)(xi[c] - 1) + (size_t)ix * (size_t)(yj[b] - 1);
                a.ro_o[q] = ro; a.pr_o[q] = pr; a.vx_o[q] = vx; a.vy_o[q] = vy;
                a.bx_o[q] = bx; a.by_o[q] = by; a.az_o[q] = az;
                if (do_cfl) {   // cans2d/common/cfl_m.f:46-55 for this cell
                    const double b2 = bx * bx + by * by;
                    const double ca2 = b2 * pi4i / ro;
                    const double cs2 = gm * pr / ro;
                    const double v2 = vx * vx + vy * vy;
                    const double dtcfl = fminF(a.dx[xi[c] - 1], a.dy[yj[b] - 1]) / __builtin_sqrt(v2 + cs2 + ca2);
                    const double dtq = a.safety * dtcfl;
                    if (dtq < dtmin_local) dtmin_local = dtq;
                }
            }
        }
    }
    if (do_cfl) block_min_to_global(dtmin_local, red, &a.clk->dtkey[(a.step + 1) % 3]);
}

}  // namespace

hipError_t launch_mlw2d(const Mlw2dArgs &a, hipStream_t s) {
    if (a.ix < 3 || a.jx < 3) return hipSuccess;
    const int ntx = (a.ix - 2 + TX - 1) / TX;
    const int nty = (a.jx - 2 + TY - 1) / TY;
    hipLaunchKernelGGL(mlw2d_kernel, dim3(ntx * nty), dim3(NT), 0, s, a);
    return hipGetLastError();
}

}  // namespace cans
