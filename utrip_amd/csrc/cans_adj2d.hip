// cans_adj2d.hip -- reverse mode (adjoint) of the fused periodic 2D MLW step
// (cans2d/hdmlw/mlw_m.f:2-324 + bdperx/bdpery + cfl_m of the next step) on
// gfx950, for gradients of a loss on the final primitives (ro, pr, vx, vy,
// bx, by) with respect to the initial state.
//
// One step forward is  P' = bnd(prim(U(P) + du(P, dt))),  dt = cfl(P).  Its
// transpose, given lam' = dL/dP' on every cell, is built as gathers (each
// output reads what it needs; no atomics on fields):
//   fold   adjoint of bdperx then bdpery: ghost lam is added to its source
//          cell (y first, then x) and cleared;
//   B1     per interior cell: lam_du = (dP'/dU')^T lam'   (prim recovery);
//   B2     per corner: recompute the half step (mlwhalf.f:45-52) and corner
//          fluxes, gather lam_du of its 4 cells through the full-step
//          weights (mlwfull.f:31-41), back through the corner flux and
//          primitive recovery: lam_uh (6 per corner);
//   B3     per cell: gather from its own update, the 4 neighbours' updates
//          (centre term mlwhalf.f:36-41, artificial viscosity
//          mlwartv.f:33-41 incl. the max(0,|dv|-1e-4) coefficients) and the
//          4 corners; back through U(P), the cell fluxes and E = -v x B;
//   dt     dL/ddt from every explicit dt and the half step, reduced on the
//          device, added to the cells whose cfl term equals dt (split evenly
//          over exact ties: the mean of the one-sided derivatives).
// az is write-only in the step (nothing reads it but its own update), so it
// carries no adjoint.  Everything is FP64.
#include "cans_common.hpp"
#include "cans_kernels.hpp"

namespace cans {
namespace {

constexpr double pi8i = 1.0 / kPi / 8.0;   // mlw_m.f:27
constexpr double pi4i = 1.0 / kPi / 4.0;   // mlw_m.f:28

struct AdjArgs {
    const double *p[6];      // P^n: ro pr vx vy bx by
    const double *q[6];      // P^{n+1}
    double *lin[6];          // lam' (folded in place)
    double *lout[6];         // lam of P^n
    double *ldu[6];          // per cell
    double *luh[6];          // per corner (ci,cj) at (ci-1)+ix*(cj-1)
    const double *dx, *dxm, *dy, *dym;
    Metric mx, my;
    double qav, gm, safety;
    int ix, jx, margin;
    const unsigned long long *key;   // dt of the step (order-preserving key)
    double *ldt;                     // dL/ddt accumulator
    unsigned *nties;
};

struct P6 {
    double ro, pr, vx, vy, bx, by;
};

CANS_DEV size_t at(int i, int j, int ix) { return (size_t)(i - 1) + (size_t)ix * (size_t)(j - 1); }

CANS_DEV P6 load6(const double *const *f, size_t k) {
    return P6{f[0][k], f[1][k], f[2][k], f[3][k], f[4][k], f[5][k]};
}

// conserved variables and fluxes of one cell, the forward's expressions
struct CellV {
    double u[6];      // ro ee rx ry bx by
    double fx[6];     // x fluxes: rx eex rxx ryx 0 -ez
    double gy[6];     // y fluxes: ry eey rxy ryy ez 0
};
CANS_DEV CellV cellv(const P6 &p, double gm) {
    CellV c;
    const double vv = p.vx * p.vx + p.vy * p.vy;
    const double bb = p.bx * p.bx + p.by * p.by;
    const double ez = -p.vx * p.by + p.vy * p.bx;
    const double ep = p.pr * gm / (gm - 1.0) + 0.5 * p.ro * vv;
    c.u[0] = p.ro;
    c.u[1] = p.pr / (gm - 1.0) + 0.5 * p.ro * vv + pi8i * bb;
    c.u[2] = p.ro * p.vx;
    c.u[3] = p.ro * p.vy;
    c.u[4] = p.bx;
    c.u[5] = p.by;
    c.fx[0] = c.u[2];
    c.fx[1] = ep * p.vx - p.by * ez * pi4i;
    c.fx[2] = p.ro * (p.vx * p.vx) + p.pr + pi8i * (p.by * p.by - p.bx * p.bx);
    c.fx[3] = p.ro * p.vy * p.vx - pi4i * p.by * p.bx;
    c.fx[4] = 0.0;
    c.fx[5] = -ez;
    c.gy[0] = c.u[3];
    c.gy[1] = ep * p.vy + p.bx * ez * pi4i;
    c.gy[2] = p.ro * p.vx * p.vy - pi4i * p.bx * p.by;
    c.gy[3] = p.ro * (p.vy * p.vy) + p.pr + pi8i * (p.bx * p.bx - p.by * p.by);
    c.gy[4] = ez;
    c.gy[5] = 0.0;
    return c;
}

CANS_DEV P6 prims(const double *u, double gm) {   // mlw_m.f:299-321
    P6 p;
    p.ro = u[0];
    p.vx = u[2] / u[0];
    p.vy = u[3] / u[0];
    const double vv = p.vx * p.vx + p.vy * p.vy;
    const double bb = u[4] * u[4] + u[5] * u[5];
    p.pr = (gm - 1.0) * (u[1] - 0.5 * u[0] * vv - pi8i * bb);
    p.bx = u[4];
    p.by = u[5];
    return p;
}

// lam of the primitives -> lam of the conserved variables they came from
CANS_DEV void prim_adj(const P6 &p, const P6 &l, double gm, double *lu) {
    const double g1 = gm - 1.0;
    const double vv = p.vx * p.vx + p.vy * p.vy;
    lu[0] = l.ro + l.pr * g1 * 0.5 * vv - (l.vx * p.vx + l.vy * p.vy) / p.ro;
    lu[1] = l.pr * g1;
    lu[2] = l.vx / p.ro - l.pr * g1 * p.vx;
    lu[3] = l.vy / p.ro - l.pr * g1 * p.vy;
    lu[4] = l.bx - l.pr * g1 * 2.0 * pi8i * p.bx;
    lu[5] = l.by - l.pr * g1 * 2.0 * pi8i * p.by;
}

// lam of U(P) -> lam of P (added)
CANS_DEV void cons_adj(const P6 &p, const double *lu, double gm, P6 &l) {
    const double vv = p.vx * p.vx + p.vy * p.vy;
    l.ro += lu[0] + lu[1] * 0.5 * vv + lu[2] * p.vx + lu[3] * p.vy;
    l.pr += lu[1] / (gm - 1.0);
    l.vx += lu[1] * p.ro * p.vx + lu[2] * p.ro;
    l.vy += lu[1] * p.ro * p.vy + lu[3] * p.ro;
    l.bx += lu[1] * 2.0 * pi8i * p.bx + lu[4];
    l.by += lu[1] * 2.0 * pi8i * p.by + lu[5];
}

// lam of the flux quantities (x: rx eex rxx ryx, y: ry eey rxy ryy, ez) -> lam of P (added)
CANS_DEV void flux_adj(const P6 &p, double gm, double lrx, double leex, double lrxx, double lryx, double lry,
                       double leey, double lrxy, double lryy, double lez, P6 &l) {
    const double ro = p.ro, pr = p.pr, vx = p.vx, vy = p.vy, bx = p.bx, by = p.by;
    (void)pr;
    const double vv = vx * vx + vy * vy;
    const double ez = -vx * by + vy * bx;
    const double ep = p.pr * gm / (gm - 1.0) + 0.5 * ro * vv;
    const double lez_t = lez - leex * by * pi4i + leey * bx * pi4i;
    const double lep = leex * vx + leey * vy;
    const double lxy = lrxy + lryx;
    l.ro += lrx * vx + lry * vy + lrxx * vx * vx + lxy * vx * vy + lryy * vy * vy + lep * 0.5 * vv;
    l.pr += lep * gm / (gm - 1.0) + lrxx + lryy;
    l.vx += lrx * ro + leex * ep + lrxx * 2.0 * ro * vx + lxy * ro * vy + lep * ro * vx - lez_t * by;
    l.vy += lry * ro + leey * ep + lxy * ro * vx + lryy * 2.0 * ro * vy + lep * ro * vy + lez_t * bx;
    l.bx += leey * ez * pi4i - lrxx * 2.0 * pi8i * bx - lxy * pi4i * by + lryy * 2.0 * pi8i * bx + lez_t * vy;
    l.by += -leex * ez * pi4i + lrxx * 2.0 * pi8i * by - lxy * pi4i * bx - lryy * 2.0 * pi8i * by - lez_t * vx;
}

CANS_DEV double step_dt(const AdjArgs &a) { return dunkey(*a.key); }

CANS_DEV bool interior(const AdjArgs &a, int i, int j) { return i >= 2 && i <= a.ix - 1 && j >= 2 && j <= a.jx - 1; }

template <int BT>
CANS_DEV void block_sum_add(double v, double *dst) {
    __shared__ double red[BT / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < BT / 64; w++) s += red[w];
        atomicAdd(dst, s);
    }
}

// ------------------------------------------------------------------ fold
// adjoint of bdpery (periodic rows): lam[src] += lam[ghost], lam[ghost] = 0
__global__ void fold_y(AdjArgs a) {
    const int m = a.margin, ix = a.ix, jx = a.jx;
    const long n = (long)ix * 2 * m;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % ix) + 1;
        const int g = (int)(t / ix);          // 0..2m-1
        const int j = g < m ? g + 1 : jx - m + (g - m) + 1;
        const int js = g < m ? j + (jx - 2 * m) : j - (jx - 2 * m);
        for (int f = 0; f < 6; f++) {
            a.lin[f][at(i, js, ix)] += a.lin[f][at(i, j, ix)];
            a.lin[f][at(i, j, ix)] = 0.0;
        }
    }
}
// adjoint of bdperx (periodic columns), after fold_y
__global__ void fold_x(AdjArgs a) {
    const int m = a.margin, ix = a.ix, jx = a.jx;
    const long n = (long)jx * 2 * m;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int j = (int)(t % jx) + 1;
        const int g = (int)(t / jx);
        const int i = g < m ? g + 1 : ix - m + (g - m) + 1;
        const int is = g < m ? i + (ix - 2 * m) : i - (ix - 2 * m);
        for (int f = 0; f < 6; f++) {
            a.lin[f][at(is, j, ix)] += a.lin[f][at(i, j, ix)];
            a.lin[f][at(i, j, ix)] = 0.0;
        }
    }
}

__global__ void adj_zero(double *ldt, unsigned *nties) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *ldt = 0.0;
        *nties = 0u;
    }
}

// ------------------------------------------------------------------ B1
__global__ __launch_bounds__(256) void adj_b1(AdjArgs a) {
    const size_t n = (size_t)a.ix * a.jx;
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = (int)(k % a.ix) + 1, j = (int)(k / a.ix) + 1;
    if (!interior(a, i, j)) {
        for (int f = 0; f < 6; f++) a.ldu[f][k] = 0.0;
        return;
    }
    const P6 pn = load6(a.q, k);
    const P6 l = load6(a.lin, k);
    double lu[6];
    prim_adj(pn, l, a.gm, lu);
    for (int f = 0; f < 6; f++) a.ldu[f][k] = lu[f];
}

// ------------------------------------------------------------------ B2
__global__ __launch_bounds__(256) void adj_b2(AdjArgs a) {
    const int ix = a.ix, jx = a.jx;
    const size_t n = (size_t)(ix - 1) * (jx - 1);
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double part = 0.0;
    if (t < n) {
        const int ci = (int)(t % (ix - 1)) + 1, cj = (int)(t / (ix - 1)) + 1;
        const double gm = a.gm, dt = step_dt(a), hdt = 0.5 * dt;
        const size_t kA = at(ci, cj, ix), kB = at(ci + 1, cj, ix), kC = at(ci, cj + 1, ix), kD = at(ci + 1, cj + 1, ix);
        const CellV A = cellv(load6(a.p, kA), gm), B = cellv(load6(a.p, kB), gm);
        const CellV Cc = cellv(load6(a.p, kC), gm), D = cellv(load6(a.p, kD), gm);
        const double dxim = a.mx.dim[ci - 1], dyim = a.my.dim[cj - 1];
        // half step (mlwhalf.f:47-50)
        double uh[6], dh[6];
        for (int q = 0; q < 6; q++) {
            const double u = 0.25 * (B.u[q] + A.u[q] + D.u[q] + Cc.u[q]);
            const double dfdx = dxim * (B.fx[q] - A.fx[q] + D.fx[q] - Cc.fx[q]) / 2.0;
            const double dfdy = dyim * (Cc.gy[q] - A.gy[q] + D.gy[q] - B.gy[q]) / 2.0;
            uh[q] = u - dt * dfdx - dt * dfdy;
            dh[q] = -dfdx - dfdy;                       // d uh / d dt
        }
        const P6 ph = prims(uh, gm);
        const CellV H = cellv(ph, gm);                  // corner fluxes (H.fx, H.gy)
        // full-step weights of this corner in its 4 cells' du (mlwfull.f:31-41)
        const bool iA = interior(a, ci, cj), iB = interior(a, ci + 1, cj);
        const bool iC = interior(a, ci, cj + 1), iD = interior(a, ci + 1, cj + 1);
        const double hA = iA ? -hdt * a.mx.di[ci - 1] * a.my.u1[cj - 1] : 0.0;
        const double hB = iB ? hdt * a.mx.di[ci] * a.my.u1[cj - 1] : 0.0;
        const double hC = iC ? -hdt * a.mx.di[ci - 1] * a.my.u0[cj] : 0.0;
        const double hD = iD ? hdt * a.mx.di[ci] * a.my.u0[cj] : 0.0;
        const double kA_ = iA ? -hdt * a.my.di[cj - 1] * a.mx.u1[ci - 1] : 0.0;
        const double kB_ = iB ? -hdt * a.my.di[cj - 1] * a.mx.u0[ci] : 0.0;
        const double kC_ = iC ? hdt * a.my.di[cj] * a.mx.u1[ci - 1] : 0.0;
        const double kD_ = iD ? hdt * a.my.di[cj] * a.mx.u0[ci] : 0.0;
        double lH[6], lK[6];
        for (int q = 0; q < 6; q++) {
            const double la = a.ldu[q][kA], lb = a.ldu[q][kB], lc = a.ldu[q][kC], ld = a.ldu[q][kD];
            lH[q] = hA * la + hB * lb + hC * lc + hD * ld;
            lK[q] = kA_ * la + kB_ * lb + kC_ * lc + kD_ * ld;
        }
        // explicit dt dependence of the corner terms: (coef/dt) * flux
        if (dt != 0.0) {
            double s = 0.0;
            for (int q = 0; q < 6; q++) s += lH[q] * H.fx[q] + lK[q] * H.gy[q];
            part += s / dt;
        }
        P6 lp{0, 0, 0, 0, 0, 0};
        flux_adj(ph, gm, lH[0], lH[1], lH[2], lH[3], lK[0], lK[1], lK[2], lK[3], -lH[5] + lK[4], lp);
        double lu[6];
        prim_adj(ph, lp, gm, lu);
        for (int q = 0; q < 6; q++) {
            a.luh[q][kA] = lu[q];
            part += lu[q] * dh[q];
        }
    }
    block_sum_add<256>(part, a.ldt);
}

// ------------------------------------------------------------------ B3
CANS_DEV double av_q(double qav, double dm, double vhi, double vlo, double &dq) {
    const double d = vhi - vlo;
    const double s = __builtin_fabs(d) - kR4_1em4;
    dq = s > 0.0 ? qav * dm * (d > 0.0 ? 1.0 : (d < 0.0 ? -1.0 : 0.0)) : 0.0;   // dq/d(vhi)
    return qav * dm * fmaxF(0.0, s);
}

__global__ __launch_bounds__(256) void adj_b3(AdjArgs a) {
    const int ix = a.ix, jx = a.jx;
    const size_t n = (size_t)ix * jx;
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double part = 0.0;
    if (k < n) {
        const int i = (int)(k % ix) + 1, j = (int)(k / ix) + 1;
        const double gm = a.gm, dt = step_dt(a), hdt = 0.5 * dt;
        const bool hasE = i < ix, hasW = i > 1, hasN = j < jx, hasS = j > 1;
        const bool ic = interior(a, i, j);
        const bool iE = hasE && interior(a, i + 1, j), iW = hasW && interior(a, i - 1, j);
        const bool iN = hasN && interior(a, i, j + 1), iS = hasS && interior(a, i, j - 1);
        const P6 pc = load6(a.p, k);
        const CellV C = cellv(pc, gm);
        const size_t kE = hasE ? k + 1 : k, kW = hasW ? k - 1 : k, kN = hasN ? k + ix : k, kS = hasS ? k - ix : k;
        const P6 pE = load6(a.p, kE), pW = load6(a.p, kW), pN = load6(a.p, kN), pS = load6(a.p, kS);
        const CellV E = cellv(pE, gm), W = cellv(pW, gm), N = cellv(pN, gm), S = cellv(pS, gm);
        double lc[6], le[6], lw[6], ln[6], ls[6];
        for (int q = 0; q < 6; q++) {
            lc[q] = ic ? a.ldu[q][k] : 0.0;
            le[q] = iE ? a.ldu[q][kE] : 0.0;
            lw[q] = iW ? a.ldu[q][kW] : 0.0;
            ln[q] = iN ? a.ldu[q][kN] : 0.0;
            ls[q] = iS ? a.ldu[q][kS] : 0.0;
        }
        // metric of this cell and its neighbours (u0/u1 are 0 on ghosts)
        const double dxi_c = a.mx.di[i - 1], dyi_c = a.my.di[j - 1];
        const double dxim_c = a.mx.dim[i - 1], dyim_c = a.my.dim[j - 1];
        const double dxim_w = hasW ? a.mx.dim[i - 2] : 0.0, dyim_s = hasS ? a.my.dim[j - 2] : 0.0;
        const double dxi_e = hasE ? a.mx.di[i] : 0.0, dxi_w = hasW ? a.mx.di[i - 2] : 0.0;
        const double dyi_n = hasN ? a.my.di[j] : 0.0, dyi_s = hasS ? a.my.di[j - 2] : 0.0;
        // AV coefficients of the four faces of this cell (mlw_m.f:276-288)
        double dqE = 0, dqW = 0, dqN = 0, dqS = 0;
        const double qE = hasE ? av_q(a.qav, a.dxm[i - 1], pE.vx, pc.vx, dqE) : 0.0;   // face (c,E)
        const double qW = hasW ? av_q(a.qav, a.dxm[i - 2], pc.vx, pW.vx, dqW) : 0.0;   // face (W,c)
        const double qN = hasN ? av_q(a.qav, a.dym[j - 1], pN.vy, pc.vy, dqN) : 0.0;   // face (c,N)
        const double qS = hasS ? av_q(a.qav, a.dym[j - 2], pc.vy, pS.vy, dqS) : 0.0;   // face (S,c)

        double lu[6], lf[6], lg[6];
        for (int q = 0; q < 6; q++) {
            // own update: identity, and the AV centre coefficient
            double u = lc[q] - lc[q] * dt * (dxi_c * (qE * dxim_c + qW * dxim_w) + dyi_c * (qN * dyim_c + qS * dyim_s));
            // neighbours' updates: their centre term reads this cell's fluxes, their AV this cell's u
            u += le[q] * dt * dxi_e * qE * dxim_c + lw[q] * dt * dxi_w * qW * dxim_w;
            u += ln[q] * dt * dyi_n * qN * dyim_c + ls[q] * dt * dyi_s * qS * dyim_s;
            double f = le[q] * hdt * (0.5 * dxi_e) - lw[q] * hdt * (0.5 * dxi_w);
            double g = ln[q] * hdt * (0.5 * dyi_n) - ls[q] * hdt * (0.5 * dyi_s);
            lu[q] = u;
            lf[q] = f;
            lg[q] = g;
        }
        // corners around this cell: (i,j) A, (i-1,j) B, (i,j-1) C, (i-1,j-1) D
        for (int r = 0; r < 4; r++) {
            const int ci = (r & 1) ? i - 1 : i, cj = (r & 2) ? j - 1 : j;
            if (ci < 1 || ci > ix - 1 || cj < 1 || cj > jx - 1) continue;
            const size_t kc = at(ci, cj, ix);
            const double sx = (r & 1) ? -1.0 : 1.0, sy = (r & 2) ? -1.0 : 1.0;
            const double wx = sx * dt * a.mx.dim[ci - 1] / 2.0, wy = sy * dt * a.my.dim[cj - 1] / 2.0;
            for (int q = 0; q < 6; q++) {
                const double l = a.luh[q][kc];
                lu[q] += 0.25 * l;
                lf[q] += wx * l;
                lg[q] += wy * l;
            }
        }
        // AV coefficient derivatives (velocities enter through max(0,|dv|-1e-4))
        double lvx = 0.0, lvy = 0.0;
        if (hasE) {   // face (c,E): in du(c) with +dt*dxi_c*q*dxim_c*(uE-uc), in du(E) with -dt*dxi_e*q*dxim_c*(uE-uc)
            double lq = 0.0;
            for (int q = 0; q < 6; q++) lq += (E.u[q] - C.u[q]) * dt * dxim_c * (dxi_c * lc[q] - dxi_e * le[q]);
            lvx -= lq * dqE;
        }
        if (hasW) {
            double lq = 0.0;
            for (int q = 0; q < 6; q++) lq += (C.u[q] - W.u[q]) * dt * dxim_w * (dxi_w * lw[q] - dxi_c * lc[q]);
            lvx += lq * dqW;
        }
        if (hasN) {
            double lq = 0.0;
            for (int q = 0; q < 6; q++) lq += (N.u[q] - C.u[q]) * dt * dyim_c * (dyi_c * lc[q] - dyi_n * ln[q]);
            lvy -= lq * dqN;
        }
        if (hasS) {
            double lq = 0.0;
            for (int q = 0; q < 6; q++) lq += (C.u[q] - S.u[q]) * dt * dyim_s * (dyi_s * ls[q] - dyi_c * lc[q]);
            lvy += lq * dqS;
        }
        P6 l{0.0, 0.0, lvx, lvy, 0.0, 0.0};
        cons_adj(pc, lu, gm, l);
        // x fluxes rx eex rxx ryx (bx: none, by: -ez); y fluxes ry eey rxy ryy (bx: ez, by: none)
        flux_adj(pc, gm, lf[0], lf[1], lf[2], lf[3], lg[0], lg[1], lg[2], lg[3], -lf[5] + lg[4], l);
        a.lout[0][k] = l.ro; a.lout[1][k] = l.pr; a.lout[2][k] = l.vx;
        a.lout[3][k] = l.vy; a.lout[4][k] = l.bx; a.lout[5][k] = l.by;
        // explicit dt dependence of this cell's update (centre term and AV)
        if (ic) {
            for (int q = 0; q < 6; q++) {
                const double d = -0.25 * dxi_c * (E.fx[q] - W.fx[q]) - 0.25 * dyi_c * (N.gy[q] - S.gy[q])
                               + dxi_c * (qE * dxim_c * (E.u[q] - C.u[q]) - qW * dxim_w * (C.u[q] - W.u[q]))
                               + dyi_c * (qN * dyim_c * (N.u[q] - C.u[q]) - qS * dyim_s * (C.u[q] - S.u[q]));
                part += lc[q] * d;
            }
        }
    }
    block_sum_add<256>(part, a.ldt);
}

// ------------------------------------------------------------------ dt = cfl(P)
CANS_DEV double cfl_term(const AdjArgs &a, int i, int j, const P6 &p, double *S) {
    const double b2 = p.bx * p.bx + p.by * p.by;
    const double ca2 = b2 * pi4i / p.ro;
    const double cs2 = a.gm * p.pr / p.ro;
    const double v2 = p.vx * p.vx + p.vy * p.vy;
    *S = v2 + cs2 + ca2;
    const double dtcfl = fminF(a.dx[i - 1], a.dy[j - 1]) / __builtin_sqrt(*S);
    return a.safety * dtcfl;
}

__global__ __launch_bounds__(256) void adj_ties(AdjArgs a) {
    const size_t n = (size_t)a.ix * a.jx;
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = (int)(k % a.ix) + 1, j = (int)(k / a.ix) + 1;
    double S;
    if (cfl_term(a, i, j, load6(a.p, k), &S) == step_dt(a)) atomicAdd(a.nties, 1u);
}

__global__ __launch_bounds__(256) void adj_dt(AdjArgs a) {
    const size_t n = (size_t)a.ix * a.jx;
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = (int)(k % a.ix) + 1, j = (int)(k / a.ix) + 1;
    const P6 p = load6(a.p, k);
    double S;
    const double q = cfl_term(a, i, j, p, &S);
    if (q != step_dt(a) || *a.nties == 0u) return;
    // d(dtq)/dS = -dtq/(2S); S = v2 + gm pr/ro + b2 pi4i/ro
    const double w = (*a.ldt / (double)(*a.nties)) * (-0.5 * q / S);
    const double b2 = p.bx * p.bx + p.by * p.by;
    a.lout[0][k] += w * (-(a.gm * p.pr + b2 * pi4i) / (p.ro * p.ro));
    a.lout[1][k] += w * (a.gm / p.ro);
    a.lout[2][k] += w * (2.0 * p.vx);
    a.lout[3][k] += w * (2.0 * p.vy);
    a.lout[4][k] += w * (2.0 * p.bx * pi4i / p.ro);
    a.lout[5][k] += w * (2.0 * p.by * pi4i / p.ro);
}

}  // namespace

hipError_t launch_adj2d_step(const double *const *pn, const double *const *pn1, double *const *lin,
                             double *const *lout, double *ws, const double *dx, const double *dxm, const double *dy,
                             const double *dym, Metric mx, Metric my, double qav, double gm, double safety, int ix,
                             int jx, int margin, hipStream_t s) {
    const size_t n = (size_t)ix * jx;
    const size_t na = (n + 63) & ~size_t(63);
    AdjArgs a{};
    for (int f = 0; f < 6; f++) {
        a.p[f] = pn[f];
        a.q[f] = pn1[f];
        a.lin[f] = lin[f];
        a.lout[f] = lout[f];
        a.ldu[f] = ws + f * na;
        a.luh[f] = ws + (6 + f) * na;
    }
    unsigned long long *key = reinterpret_cast<unsigned long long *>(ws + 12 * na);
    a.ldt = ws + 12 * na + 8;
    a.nties = reinterpret_cast<unsigned *>(ws + 12 * na + 16);
    a.key = key;
    a.dx = dx; a.dxm = dxm; a.dy = dy; a.dym = dym;
    a.mx = mx; a.my = my;
    a.qav = qav; a.gm = gm; a.safety = safety;
    a.ix = ix; a.jx = jx; a.margin = margin;
    const int nb = (int)((n + 255) / 256);
    hipError_t e;
    hipLaunchKernelGGL(fold_y, dim3(64), dim3(256), 0, s, a);
    hipLaunchKernelGGL(fold_x, dim3(64), dim3(256), 0, s, a);
    hipLaunchKernelGGL(adj_zero, dim3(1), dim3(64), 0, s, a.ldt, a.nties);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // dt of the step: cfl_m over every cell of P^n (the forward's fused reduction)
    if ((e = launch_arm_key(key, s)) != hipSuccess) return e;
    CflArgs c{};
    c.ro = pn[0]; c.pr = pn[1]; c.vx = pn[2]; c.vy = pn[3]; c.bx = pn[4]; c.by = pn[5];
    c.dx = dx; c.dy = dy; c.safety = safety; c.gm = gm; c.ix = ix; c.jx = jx; c.kx = 1; c.dim = 2;
    c.key = key;
    if ((e = launch_cfl(c, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(adj_b1, dim3(nb), dim3(256), 0, s, a);
    hipLaunchKernelGGL(adj_b2, dim3((int)(((size_t)(ix - 1) * (jx - 1) + 255) / 256)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(adj_b3, dim3(nb), dim3(256), 0, s, a);
    hipLaunchKernelGGL(adj_ties, dim3(nb), dim3(256), 0, s, a);
    hipLaunchKernelGGL(adj_dt, dim3(nb), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_periodic_fold(double *const *lam, int ix, int jx, int margin, hipStream_t s) {
    AdjArgs a{};
    for (int f = 0; f < 6; f++) a.lin[f] = lam[f];
    a.ix = ix; a.jx = jx; a.margin = margin;
    hipLaunchKernelGGL(fold_y, dim3(64), dim3(256), 0, s, a);
    hipLaunchKernelGGL(fold_x, dim3(64), dim3(256), 0, s, a);
    return hipGetLastError();
}

__global__ void copy6_kernel(AdjArgs a, size_t n) {
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        for (int f = 0; f < 6; f++) a.lout[f][k] = a.lin[f][k];
}

hipError_t launch_copy6(double *const *dst, double *const *src, size_t n, hipStream_t s) {
    AdjArgs a{};
    for (int f = 0; f < 6; f++) {
        a.lout[f] = dst[f];
        a.lin[f] = src[f];
    }
    hipLaunchKernelGGL(copy6_kernel, dim3(512), dim3(256), 0, s, a, n);
    return hipGetLastError();
}

size_t adj2d_ws_doubles(int ix, int jx) {
    const size_t n = (size_t)ix * jx;
    const size_t na = (n + 63) & ~size_t(63);
    return 12 * na + 64;
}

}  // namespace cans
