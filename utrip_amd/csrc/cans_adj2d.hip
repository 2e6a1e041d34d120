// cans_adj2d.hip -- reverse mode (adjoint) of the fused periodic 2D MLW step
// (cans2d/hdmlw/mlw_m.f:2-324 + bdperx/bdpery + cfl_m of the next step) on
// gfx950, for gradients of a loss on the final primitives (ro, pr, vx, vy,
// bx, by) with respect to the initial state.
//
// One step forward is  P' = bnd(prim(U(P) + du(P, dt))),  dt = cfl(P).  Its
// transpose, given lam' = dL/dP' on every cell, is built as gathers (each
// output reads what it needs; no atomics on fields):
//   B1     per interior cell: the adjoint of the periodic fill (the cell's
//          lam' plus that of its ghost images, the images bdperx then
//          bdpery made of it), then lam_du = (dP'/dU')^T lam' (primitive
//          recovery); every cell also counts whether its cfl term ties dt;
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
//          over exact ties: the mean of the one-sided derivatives).  The dt
//          of each step is the one the forward kernel logged.
// az is write-only in the step (nothing reads it but its own update), so it
// carries no adjoint.  Everything is FP64.
#include "cans_common.hpp"
#include "cans_kernels.hpp"
#include "cans_roe.hpp"
#include "cans_roeflux_adj.hpp"

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
    const double *dtp;               // the dt of the step
    double *ldt;                     // dL/ddt accumulator stripes (zeroed)
    unsigned *nties;                 // cells whose cfl term equals dt (zeroed)
    const double *pend_dt;           // the dt of step n+1, whose dt gradient B1 applies (or NULL)
    const double *pend_acc;          // its accumulator slot (ldt stripes, nties)
    double *acc_zero;                // accumulators B2 clears for step n-1 (or NULL)
    int ilo;                         // interior() lower bound
};

struct P6 {
    double ro, pr, vx, vy, bx, by;
};

// The reverse kernels take their blocks in the XCD-contiguous order of
// xcd_first_thread() (cans_common.hpp): the neighbour rows a cell's gathers
// read are then mostly fetched into one L2.
CANS_DEV size_t at(int i, int j, int ix) { return (size_t)(i - 1) + (size_t)ix * (size_t)(j - 1); }

CANS_DEV P6 load6(const double *const *f, size_t k) {
    return P6{f[0][k], f[1][k], f[2][k], f[3][k], f[4][k], f[5][k]};
}

// The reverse sweep re-evaluates the forward's cell values, half step and
// primitives only as the point its Jacobians are taken at: it divides by
// gm-1 and by rho through reciprocals (g1i = 1/(gm-1), one 1/rho per
// cell), which moves that point by at most an ulp and the gradient by
// ~1e-16 relative.  The one forward value it must match bit for bit is the
// cfl term, compared with dt for the ties (cfl_term, unchanged).

// conserved variables and fluxes of one cell (the forward's expressions)
struct CellV {
    double u[6];      // ro ee rx ry bx by
    double fx[6];     // x fluxes: rx eex rxx ryx 0 -ez
    double gy[6];     // y fluxes: ry eey rxy ryy ez 0
};
CANS_DEV CellV cellv(const P6 &p, double gm, double g1i) {
    CellV c;
    const double vv = p.vx * p.vx + p.vy * p.vy;
    const double bb = p.bx * p.bx + p.by * p.by;
    const double ez = -p.vx * p.by + p.vy * p.bx;
    const double ep = p.pr * gm * g1i + 0.5 * p.ro * vv;
    c.u[0] = p.ro;
    c.u[1] = p.pr * g1i + 0.5 * p.ro * vv + pi8i * bb;
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
    const double rr = 1.0 / u[0];
    p.vx = u[2] * rr;
    p.vy = u[3] * rr;
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
    const double rr = 1.0 / p.ro;
    lu[0] = l.ro + l.pr * g1 * 0.5 * vv - (l.vx * p.vx + l.vy * p.vy) * rr;
    lu[1] = l.pr * g1;
    lu[2] = l.vx * rr - l.pr * g1 * p.vx;
    lu[3] = l.vy * rr - l.pr * g1 * p.vy;
    lu[4] = l.bx - l.pr * g1 * 2.0 * pi8i * p.bx;
    lu[5] = l.by - l.pr * g1 * 2.0 * pi8i * p.by;
}

// lam of U(P) -> lam of P (added)
CANS_DEV void cons_adj(const P6 &p, const double *lu, double g1i, P6 &l) {
    const double vv = p.vx * p.vx + p.vy * p.vy;
    l.ro += lu[0] + lu[1] * 0.5 * vv + lu[2] * p.vx + lu[3] * p.vy;
    l.pr += lu[1] * g1i;
    l.vx += lu[1] * p.ro * p.vx + lu[2] * p.ro;
    l.vy += lu[1] * p.ro * p.vy + lu[3] * p.ro;
    l.bx += lu[1] * 2.0 * pi8i * p.bx + lu[4];
    l.by += lu[1] * 2.0 * pi8i * p.by + lu[5];
}

// lam of the flux quantities (x: rx eex rxx ryx, y: ry eey rxy ryy, ez) -> lam of P (added)
CANS_DEV void flux_adj(const P6 &p, double gm, double g1i, double lrx, double leex, double lrxx, double lryx, double lry,
                       double leey, double lrxy, double lryy, double lez, P6 &l) {
    const double ro = p.ro, pr = p.pr, vx = p.vx, vy = p.vy, bx = p.bx, by = p.by;
    (void)pr;
    const double vv = vx * vx + vy * vy;
    const double ez = -vx * by + vy * bx;
    const double ep = p.pr * gm * g1i + 0.5 * ro * vv;
    const double lez_t = lez - leex * by * pi4i + leey * bx * pi4i;
    const double lep = leex * vx + leey * vy;
    const double lxy = lrxy + lryx;
    l.ro += lrx * vx + lry * vy + lrxx * vx * vx + lxy * vx * vy + lryy * vy * vy + lep * 0.5 * vv;
    l.pr += lep * gm * g1i + lrxx + lryy;
    l.vx += lrx * ro + leex * ep + lrxx * 2.0 * ro * vx + lxy * ro * vy + lep * ro * vx - lez_t * by;
    l.vy += lry * ro + leey * ep + lxy * ro * vx + lryy * 2.0 * ro * vy + lep * ro * vy + lez_t * bx;
    l.bx += leey * ez * pi4i - lrxx * 2.0 * pi8i * bx - lxy * pi4i * by + lryy * 2.0 * pi8i * bx + lez_t * vy;
    l.by += -leex * ez * pi4i + lrxx * 2.0 * pi8i * by - lxy * pi4i * bx - lryy * 2.0 * pi8i * by - lez_t * vx;
}

CANS_DEV double step_dt(const AdjArgs &a) { return *a.dtp; }

// cells whose update survives the periodic fill: [ilo, n+1-ilo] (MLW: ilo = 2,
// the engine's update range; Roe: margin+1, the interior of the box)
CANS_DEV bool interior(const AdjArgs &a, int i, int j) {
    return i >= a.ilo && i <= a.ix + 1 - a.ilo && j >= a.ilo && j <= a.jx + 1 - a.ilo;
}

// the dL/ddt of a slot: its stripes summed in order
CANS_DEV double acc_ldt(const double *acc) {
    double s = acc[0];
    for (int k = 1; k < kAdjStripes; k++) s += acc[k];
    return s;
}

// block sum added to one of the kAdjStripes stripes at dst
template <int BT>
CANS_DEV void block_sum_add(double v, double *dst) {
    __shared__ double red[BT / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < BT / 64; w++) s += red[w];
        atomicAdd(dst + (blockIdx.x % kAdjStripes), s);
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

// dL/dP at a cell whose cfl term ties dt: (dL/ddt / nties) d(dtq)/dP, with
// d(dtq)/dS = -dtq/(2S), S = v2 + gm pr/ro + b2 pi4i/ro
CANS_DEV void dt_grad(const P6 &p, double q, double S, double ldt, unsigned nties, double gm, P6 &l) {
    const double w = (ldt / (double)nties) * (-0.5 * q / S);
    const double b2 = p.bx * p.bx + p.by * p.by;
    l.ro += w * (-(gm * p.pr + b2 * pi4i) / (p.ro * p.ro));
    l.pr += w * (gm / p.ro);
    l.vx += w * (2.0 * p.vx);
    l.vy += w * (2.0 * p.vy);
    l.bx += w * (2.0 * p.bx * pi4i / p.ro);
    l.by += w * (2.0 * p.by * pi4i / p.ro);
}

// ------------------------------------------------------------------ B1
__global__ __launch_bounds__(256) void adj_b1(AdjArgs a) {
    const int ix = a.ix, jx = a.jx;
    const size_t n = (size_t)ix * jx;
    const size_t k = xcd_first_thread() + threadIdx.x;
    if (k >= n) return;
    const int i = (int)(k % ix) + 1, j = (int)(k / ix) + 1;
    double S;
    if (cfl_term(a, i, j, load6(a.p, k), &S) == step_dt(a)) atomicAdd(a.nties, 1u);
    if (!interior(a, i, j)) {
        for (int f = 0; f < 6; f++) a.ldu[f][k] = 0.0;
        return;
    }
    // adjoint of the periodic fill: after bdperx then bdpery ghost (gi,gj)
    // holds the interior (src_x(gi), src_y(gj)), so this cell collects the
    // lam' of every (image column, image row) pair it was copied to
    const int m = a.margin;
    int xi[3] = {i, 0, 0}, yj[3] = {j, 0, 0};
    int nxi = 1, nyj = 1;
    if (i >= ix - 2 * m + 1) xi[nxi++] = i - (ix - 2 * m);
    if (i <= 2 * m) xi[nxi++] = i + (ix - 2 * m);
    if (j >= jx - 2 * m + 1) yj[nyj++] = j - (jx - 2 * m);
    if (j <= 2 * m) yj[nyj++] = j + (jx - 2 * m);
    const P6 pn = load6(a.q, k);
    // the pending dt gradient of step n+1 (dt_{n+1} = cfl_m(P^{n+1}) over every
    // cell, ghost images included): added at the images that tie
    const bool pend = a.pend_dt != nullptr;
    const double pdt = pend ? *a.pend_dt : 0.0;
    const double pldt = pend ? acc_ldt(a.pend_acc) : 0.0;
    const unsigned pnt = pend ? *reinterpret_cast<const unsigned *>(a.pend_acc + kAdjStripes) : 0u;
    P6 l{0, 0, 0, 0, 0, 0};
    for (int b = 0; b < nyj; b++)
        for (int c = 0; c < nxi; c++) {
            const P6 g = load6(a.lin, at(xi[c], yj[b], ix));
            l.ro += g.ro; l.pr += g.pr; l.vx += g.vx; l.vy += g.vy; l.bx += g.bx; l.by += g.by;
            if (pnt) {
                double Sg;
                const double qg = cfl_term(a, xi[c], yj[b], pn, &Sg);
                if (qg == pdt) dt_grad(pn, qg, Sg, pldt, pnt, a.gm, l);
            }
        }
    double lu[6];
    prim_adj(pn, l, a.gm, lu);
    for (int f = 0; f < 6; f++) a.ldu[f][k] = lu[f];
}

// ------------------------------------------------------------------ B2
__global__ __launch_bounds__(256) void adj_b2(AdjArgs a) {
    const int ix = a.ix, jx = a.jx;
    const size_t n = (size_t)(ix - 1) * (jx - 1);
    const size_t t = xcd_first_thread() + threadIdx.x;
    double part = 0.0;
    if (t < n) {
        const int ci = (int)(t % (ix - 1)) + 1, cj = (int)(t / (ix - 1)) + 1;
        const double gm = a.gm, g1i = 1.0 / (gm - 1.0), dt = step_dt(a), hdt = 0.5 * dt;
        if (t == 0 && a.acc_zero) {   // B1 of this step has read them; B1 of step n-1 counts into them
            for (int q = 0; q < kAdjAccSlot; q++) a.acc_zero[q] = 0.0;
        }
        const size_t kA = at(ci, cj, ix), kB = at(ci + 1, cj, ix), kC = at(ci, cj + 1, ix), kD = at(ci + 1, cj + 1, ix);
        const CellV A = cellv(load6(a.p, kA), gm, g1i), B = cellv(load6(a.p, kB), gm, g1i);
        const CellV Cc = cellv(load6(a.p, kC), gm, g1i), D = cellv(load6(a.p, kD), gm, g1i);
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
        const CellV H = cellv(ph, gm, g1i);             // corner fluxes (H.fx, H.gy)
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
        flux_adj(ph, gm, g1i, lH[0], lH[1], lH[2], lH[3], lK[0], lK[1], lK[2], lK[3], -lH[5] + lK[4], lp);
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

// One neighbour's share of this cell's gather.  nb: its primitives; lnb:
// its lam_du (zero unless it is interior); dir 0 E, 1 W, 2 N, 3 S.
struct Gather {
    double lu[6], lf[6], lg[6];
    double lvx, lvy, part;
};

__global__ __launch_bounds__(256) void adj_b3(AdjArgs a) {
    const int ix = a.ix, jx = a.jx;
    const size_t n = (size_t)ix * jx;
    const size_t k = xcd_first_thread() + threadIdx.x;
    double part = 0.0;
    if (k < n) {
        const int i = (int)(k % ix) + 1, j = (int)(k / ix) + 1;
        const double gm = a.gm, g1i = 1.0 / (gm - 1.0), dt = step_dt(a), hdt = 0.5 * dt;
        const bool ic = interior(a, i, j);
        const P6 pc = load6(a.p, k);
        const CellV C = cellv(pc, gm, g1i);
        double lc[6];
        for (int q = 0; q < 6; q++) lc[q] = ic ? a.ldu[q][k] : 0.0;
        const double dxi_c = a.mx.di[i - 1], dyi_c = a.my.di[j - 1];
        const double dxim_c = a.mx.dim[i - 1], dyim_c = a.my.dim[j - 1];
        Gather G;
        for (int q = 0; q < 6; q++) {
            G.lu[q] = lc[q];          // own update: U' = U + du
            G.lf[q] = 0.0;
            G.lg[q] = 0.0;
        }
        G.lvx = 0.0;
        G.lvy = 0.0;
        double avc = 0.0;             // own AV centre coefficient (sum over the 4 faces)
        // the four neighbours, one at a time
        for (int d = 0; d < 4; d++) {
            const bool x = d < 2;
            const bool hi = (d == 0 || d == 2);           // E or N
            const bool has = d == 0 ? i < ix : d == 1 ? i > 1 : d == 2 ? j < jx : j > 1;
            if (!has) continue;
            const int in_ = d == 0 ? i + 1 : d == 1 ? i - 1 : i, jn = d == 2 ? j + 1 : d == 3 ? j - 1 : j;
            const size_t kn = at(in_, jn, ix);
            const bool inb = interior(a, in_, jn);
            const P6 pn = load6(a.p, kn);
            const CellV Nb = cellv(pn, gm, g1i);
            double ln[6];
            for (int q = 0; q < 6; q++) ln[q] = inb ? a.ldu[q][kn] : 0.0;
            // the face between this cell and the neighbour: its AV coefficient
            // (the face's low-side cell carries dxm; mlw_m.f:276-288)
            const int lo_i = (x && !hi) ? in_ : i, lo_j = (!x && !hi) ? jn : j;
            const double dm = x ? a.dxm[lo_i - 1] : a.dym[lo_j - 1];
            const double dmi = x ? a.mx.dim[lo_i - 1] : a.my.dim[lo_j - 1];
            const double vhi = x ? (hi ? pn.vx : pc.vx) : (hi ? pn.vy : pc.vy);
            const double vlo = x ? (hi ? pc.vx : pn.vx) : (hi ? pc.vy : pn.vy);
            double dq;
            const double qf = av_q(a.qav, dm, vhi, vlo, dq);
            const double di_n = x ? a.mx.di[in_ - 1] : a.my.di[jn - 1];
            const double di_c = x ? dxi_c : dyi_c;
            avc += di_c * qf * dmi;
            double lq = 0.0;
            for (int q = 0; q < 6; q++) {
                // neighbour's AV reads this cell's u; its centre term this cell's flux
                G.lu[q] += ln[q] * dt * di_n * qf * dmi;
                const double w = ln[q] * hdt * (0.5 * di_n);
                if (x) G.lf[q] += hi ? w : -w;
                else G.lg[q] += hi ? w : -w;
                // the face coefficient's weight: +dt*di_lo*q*dmi*(u_hi-u_lo) in du(lo), minus in du(hi)
                const double uhi = hi ? Nb.u[q] : C.u[q], ulo = hi ? C.u[q] : Nb.u[q];
                const double llo = hi ? lc[q] : ln[q], lhi = hi ? ln[q] : lc[q];
                const double dlo = hi ? di_c : di_n, dhi = hi ? di_n : di_c;
                lq += (uhi - ulo) * dt * dmi * (dlo * llo - dhi * lhi);
                // explicit dt of this cell's update: centre term and AV
                if (ic) {
                    const double fl = x ? Nb.fx[q] : Nb.gy[q];
                    part += lc[q] * ((hi ? -0.25 : 0.25) * di_c * fl
                                     + (hi ? 1.0 : -1.0) * di_c * qf * dmi * (hi ? (Nb.u[q] - C.u[q]) : (C.u[q] - Nb.u[q])));
                }
            }
            // dq is d q / d v_hi; v_hi is this cell's velocity when the neighbour is the low side
            if (x) G.lvx += hi ? -lq * dq : lq * dq;
            else G.lvy += hi ? -lq * dq : lq * dq;
        }
        for (int q = 0; q < 6; q++) G.lu[q] -= lc[q] * dt * avc;
        // corners around this cell: (i,j) A, (i-1,j) B, (i,j-1) C, (i-1,j-1) D
        for (int r = 0; r < 4; r++) {
            const int ci = (r & 1) ? i - 1 : i, cj = (r & 2) ? j - 1 : j;
            if (ci < 1 || ci > ix - 1 || cj < 1 || cj > jx - 1) continue;
            const size_t kc = at(ci, cj, ix);
            const double sx = (r & 1) ? -1.0 : 1.0, sy = (r & 2) ? -1.0 : 1.0;
            const double wx = sx * dt * a.mx.dim[ci - 1] / 2.0, wy = sy * dt * a.my.dim[cj - 1] / 2.0;
            for (int q = 0; q < 6; q++) {
                const double l = a.luh[q][kc];
                G.lu[q] += 0.25 * l;
                G.lf[q] += wx * l;
                G.lg[q] += wy * l;
            }
        }
        P6 l{0.0, 0.0, G.lvx, G.lvy, 0.0, 0.0};
        cons_adj(pc, G.lu, g1i, l);
        // x fluxes rx eex rxx ryx (bx: none, by: -ez); y fluxes ry eey rxy ryy (bx: ez, by: none)
        flux_adj(pc, gm, g1i, G.lf[0], G.lf[1], G.lf[2], G.lf[3], G.lg[0], G.lg[1], G.lg[2], G.lg[3], -G.lf[5] + G.lg[4], l);
        a.lout[0][k] = l.ro; a.lout[1][k] = l.pr; a.lout[2][k] = l.vx;
        a.lout[3][k] = l.vy; a.lout[4][k] = l.bx; a.lout[5][k] = l.by;
    }
    block_sum_add<256>(part, a.ldt);
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
    P6 l{0, 0, 0, 0, 0, 0};
    dt_grad(p, q, S, acc_ldt(a.ldt), *a.nties, a.gm, l);
    a.lout[0][k] += l.ro; a.lout[1][k] += l.pr; a.lout[2][k] += l.vx;
    a.lout[3][k] += l.vy; a.lout[4][k] += l.bx; a.lout[5][k] += l.by;
}

// ================================================================== Roe
// Reverse mode of one periodic 2D Roe-MUSCL step (cans2d/hdroe/roe_m.f:2-208
// with bdperx/bdpery and cfl_m), FP64.  Forward, per step:
//   F1 = roeflux(P_c, P_c+1) on every face (first order);
//   Uh = U(P) + 1/2 dt sum_d (F1(c-1) - F1(c))/d  on cells [2,n-1]^2,
//   Ph = prim(Uh) there and P elsewhere;
//   F2 = roeflux(L, R), L/R the minmod (tvdminmod.f) states of Ph;
//   U' = U(P) + dt sum_d (F2(c-1) - F2(c))/d  on cells [3,n-2]^2, P' = prim(U');
//   periodic fill (x then y) of margin m >= 3.
// Reverse, as gathers (no atomics on fields): B1 (adj_b1 with interior
// [m+1, n-m]: periodic fold, pending dt gradient, lam U'); R2 per stage-2
// face: lam F2 from the two cells' lam U', back through the flux (the
// generated roeflux12_vjp) and the minmod reconstruction to its four cells;
// R3 per cell: gather lam Ph, back through prim(Uh) and U(P); R4 per
// stage-1 face: lam F1 from lam Uh, back through the flux; R5 per cell:
// gather.  The face fluxes' explicit dt terms go to dL/ddt.
struct RoeAdj {
    const double *p[6];      // P^n: ro pr vx vy bx by
    const double *ph[6];     // half-step primitives (recomputed from P^n)
    double *lout[6];         // lam of P^n
    double *w;               // scratch base (SoA, stride na)
    size_t na;
    const double *dx, *dy;
    const double *dtp;
    double *ldt;
    double *acc_zero;
    double gm;
    int ix, jx;
};
enum { RW_LDU = 0, RW_LUH = 6, RW_C2 = 12, RW_C1 = 60, RW_PH = 84, RW_F1 = 92, RW_N = 108 };
CANS_DEV double *rw(const RoeAdj &a, int slot) { return a.w + (size_t)slot * a.na; }

// sweep frame of direction dir: (vn, vt, bn, bt) as field indices (ro 0,
// pr 1, vx 2, vy 3, bx 4, by 5); the flux's momentum/field outputs go to
// U slots (rn, rt, bt)
CANS_DEV void roe_frame(int dir, int &vn, int &vt, int &bn, int &bt) {
    vn = dir == 0 ? 2 : 3;
    vt = dir == 0 ? 3 : 2;
    bn = dir == 0 ? 4 : 5;
    bt = dir == 0 ? 5 : 4;
}

// d flmt(a, b) / d(a, b) for flmt = max(0, min(b sign a, |a|)) sign a
// (tvdminmod.f statement function), as evaluated: the branch taken
CANS_DEV void flmt_d(double a, double b, double &ga, double &gb) {
    const double s = signF(1.0, a);
    const double x = b * s, y = __builtin_fabs(a);
    const double m = x < y ? x : y;
    if (0.0 > m) { ga = 0.0; gb = 0.0; return; }
    if (x < y) { ga = 0.0; gb = 1.0; } else { ga = 1.0; gb = 0.0; }
}

// stage-2 faces of direction dir: x faces i in [2, ix-2], j in [3, jx-2]
// (y: transposed), between cells c and c+1 along dir.  The direction is a
// template argument: the flux slots are indexed by it, and a run-time index
// would put those small arrays in scratch.
template <int DIR>
__global__ __launch_bounds__(256) void roe_r2(RoeAdj a) {
    constexpr int dir = DIR;
    const int ix = a.ix, jx = a.jx;
    const int na_ = dir == 0 ? ix - 3 : ix - 4, nb_ = dir == 0 ? jx - 4 : jx - 3;
    const size_t t = xcd_first_thread() + threadIdx.x;
    double part = 0.0;
    if (na_ > 0 && nb_ > 0 && t < (size_t)na_ * nb_) {
        if (t == 0 && a.acc_zero && dir == 0)
            for (int q = 0; q < kAdjAccSlot; q++) a.acc_zero[q] = 0.0;
        const int i = (dir == 0 ? 2 : 3) + (int)(t % na_), j = (dir == 0 ? 3 : 2) + (int)(t / na_);
        const size_t s = dir == 0 ? 1 : (size_t)ix;
        const size_t p = at(i, j, ix);
        const double dt = *a.dtp;
        const double dn0 = dir == 0 ? a.dx[i - 1] : a.dy[j - 1], dn1 = dir == 0 ? a.dx[i] : a.dy[j];
        int vn, vt, bn, bt;
        roe_frame(dir, vn, vt, bn, bt);
        // lam F2 in flux order (ro, ee, rn, rt, bt): U'(c) gets -dt F/dn0, U'(c+1) +dt F/dn1
        const int uq[5] = {0, 1, vn, vt, bt};
        double lF[5];
        for (int k = 0; k < 5; k++) lF[k] = dt * (rw(a, RW_LDU + uq[k])[p + s] / dn1 - rw(a, RW_LDU + uq[k])[p] / dn0);
        // minmod states of the six fields from Ph at c-1 .. c+2
        double L[6], R[6], gaL[6], gbL[6], gaR[6], gbR[6];
        for (int f = 0; f < 6; f++) {
            const double *d = a.ph[f];
            const double d0 = d[p - s], d1 = d[p], d2 = d[p + s], d3 = d[p + 2 * s];
            L[f] = d1 + 0.5 * flmt(d2 - d1, d1 - d0);
            R[f] = d2 - 0.5 * flmt(d2 - d1, d3 - d2);
            flmt_d(d2 - d1, d1 - d0, gaL[f], gbL[f]);
            flmt_d(d2 - d1, d3 - d2, gaR[f], gbR[f]);
        }
        double li[12], fo[5];
        roeflux12_vjp(a.gm, L[0], L[1], L[vn], L[vt], L[bn], L[bt], R[0], R[1], R[vn], R[vt], R[bn], R[bt], lF, li, fo);
        if (dt != 0.0) {
            double sdt = 0.0;
            for (int k = 0; k < 5; k++) sdt += fo[k] * lF[k];
            part += sdt / dt;
        }
        const int fidx[6] = {0, 1, vn, vt, bn, bt};
        double lL[6], lR[6];
        for (int k = 0; k < 6; k++) {
            lL[fidx[k]] = li[k];
            lR[fidx[k]] = li[6 + k];
        }
        // contributions to cells c-1, c, c+1, c+2 (slots RW_C2 + dir*24 + off*6 + field)
        for (int f = 0; f < 6; f++) {
            const double c0 = lL[f] * (-0.5 * gbL[f]);
            const double c1 = lL[f] * (1.0 + 0.5 * (gbL[f] - gaL[f])) + lR[f] * (0.5 * gaR[f]);
            const double c2 = lL[f] * (0.5 * gaL[f]) + lR[f] * (1.0 - 0.5 * (gaR[f] - gbR[f]));
            const double c3 = lR[f] * (-0.5 * gbR[f]);
            rw(a, RW_C2 + dir * 24 + 0 * 6 + f)[p] = c0;
            rw(a, RW_C2 + dir * 24 + 1 * 6 + f)[p] = c1;
            rw(a, RW_C2 + dir * 24 + 2 * 6 + f)[p] = c2;
            rw(a, RW_C2 + dir * 24 + 3 * 6 + f)[p] = c3;
        }
    }
    block_sum_add<256>(part, a.ldt);
}

CANS_DEV bool r2_face(int dir, int i, int j, int ix, int jx) {   // stage-2 face exists
    return dir == 0 ? (i >= 2 && i <= ix - 2 && j >= 3 && j <= jx - 2) : (i >= 3 && i <= ix - 2 && j >= 2 && j <= jx - 2);
}
CANS_DEV bool r4_face(int dir, int i, int j, int ix, int jx) {   // stage-1 face exists
    return dir == 0 ? (i >= 1 && i <= ix - 1 && j >= 2 && j <= jx - 1) : (i >= 2 && i <= ix - 1 && j >= 1 && j <= jx - 1);
}

// cells: lam Ph gathered from the stage-2 faces, back through prim(Uh)
// (cells [2,n-1]^2) or the copy Ph = P (outer cells), and U(P) of the half
// step and of the update
__global__ __launch_bounds__(256) void roe_r3(RoeAdj a) {
    const int ix = a.ix, jx = a.jx;
    const size_t n = (size_t)ix * jx;
    const size_t k = xcd_first_thread() + threadIdx.x;
    if (k >= n) return;
    const int i = (int)(k % ix) + 1, j = (int)(k / ix) + 1;
    double lph[6] = {0, 0, 0, 0, 0, 0};
    for (int dir = 0; dir < 2; dir++) {
        for (int off = 0; off < 4; off++) {   // face f = c + 1 - off reads this cell as its (off)th cell
            const int fi = dir == 0 ? i + 1 - off : i, fj = dir == 0 ? j : j + 1 - off;
            if (!r2_face(dir, fi, fj, ix, jx)) continue;
            const size_t pf = at(fi, fj, ix);
            for (int f = 0; f < 6; f++) lph[f] += rw(a, RW_C2 + dir * 24 + off * 6 + f)[pf];
        }
    }
    const P6 pc = load6(a.p, k);
    double lu[6];
    for (int f = 0; f < 6; f++) lu[f] = rw(a, RW_LDU + f)[k];   // the update's U(P)
    P6 l{0, 0, 0, 0, 0, 0};
    const bool inner = i >= 2 && i <= ix - 1 && j >= 2 && j <= jx - 1;
    if (inner) {
        const P6 ph = load6(a.ph, k);
        const P6 lp{lph[0], lph[1], lph[2], lph[3], lph[4], lph[5]};
        double luh[6];
        prim_adj(ph, lp, a.gm, luh);
        for (int f = 0; f < 6; f++) {
            rw(a, RW_LUH + f)[k] = luh[f];
            lu[f] += luh[f];
        }
    } else {
        for (int f = 0; f < 6; f++) rw(a, RW_LUH + f)[k] = 0.0;
        l.ro = lph[0]; l.pr = lph[1]; l.vx = lph[2]; l.vy = lph[3]; l.bx = lph[4]; l.by = lph[5];
    }
    cons_adj(pc, lu, 1.0 / (a.gm - 1.0), l);
    a.lout[0][k] = l.ro; a.lout[1][k] = l.pr; a.lout[2][k] = l.vx;
    a.lout[3][k] = l.vy; a.lout[4][k] = l.bx; a.lout[5][k] = l.by;
}

// stage-1 faces: lam F1 from the two cells' lam Uh, back through the flux at
// (P_c, P_c+1); lam of both states stored per face
template <int DIR>
__global__ __launch_bounds__(256) void roe_r4(RoeAdj a) {
    constexpr int dir = DIR;
    const int ix = a.ix, jx = a.jx;
    const int na_ = dir == 0 ? ix - 1 : ix - 2, nb_ = dir == 0 ? jx - 2 : jx - 1;
    const size_t t = xcd_first_thread() + threadIdx.x;
    double part = 0.0;
    if (t < (size_t)na_ * nb_) {
        const int i = (dir == 0 ? 1 : 2) + (int)(t % na_), j = (dir == 0 ? 2 : 1) + (int)(t / na_);
        const size_t s = dir == 0 ? 1 : (size_t)ix;
        const size_t p = at(i, j, ix);
        const double dt = *a.dtp;
        const double dn0 = dir == 0 ? a.dx[i - 1] : a.dy[j - 1], dn1 = dir == 0 ? a.dx[i] : a.dy[j];
        int vn, vt, bn, bt;
        roe_frame(dir, vn, vt, bn, bt);
        const int uq[5] = {0, 1, vn, vt, bt};
        double lF[5];
        for (int k = 0; k < 5; k++)
            lF[k] = 0.5 * dt * (rw(a, RW_LUH + uq[k])[p + s] / dn1 - rw(a, RW_LUH + uq[k])[p] / dn0);
        const P6 pl = load6(a.p, p), pr = load6(a.p, p + s);
        const double L[6] = {pl.ro, pl.pr, pl.vx, pl.vy, pl.bx, pl.by};
        const double R[6] = {pr.ro, pr.pr, pr.vx, pr.vy, pr.bx, pr.by};
        double li[12], fo[5];
        roeflux12_vjp(a.gm, L[0], L[1], L[vn], L[vt], L[bn], L[bt], R[0], R[1], R[vn], R[vt], R[bn], R[bt], lF, li, fo);
        if (dt != 0.0) {
            double sdt = 0.0;
            for (int k = 0; k < 5; k++) sdt += fo[k] * lF[k];
            part += sdt / dt;
        }
        const int fidx[6] = {0, 1, vn, vt, bn, bt};
        for (int k = 0; k < 6; k++) {
            rw(a, RW_C1 + dir * 12 + fidx[k])[p] = li[k];
            rw(a, RW_C1 + dir * 12 + 6 + fidx[k])[p] = li[6 + k];
        }
    }
    block_sum_add<256>(part, a.ldt);
}

// cells: add the stage-1 faces' lam of this cell's state
__global__ __launch_bounds__(256) void roe_r5(RoeAdj a) {
    const int ix = a.ix, jx = a.jx;
    const size_t n = (size_t)ix * jx;
    const size_t k = xcd_first_thread() + threadIdx.x;
    if (k >= n) return;
    const int i = (int)(k % ix) + 1, j = (int)(k / ix) + 1;
    double l[6];
    for (int f = 0; f < 6; f++) l[f] = a.lout[f][k];
    for (int dir = 0; dir < 2; dir++) {
        const int li_ = dir == 0 ? i - 1 : i, lj_ = dir == 0 ? j : j - 1;   // face before: this cell is its R
        if (r4_face(dir, li_, lj_, ix, jx)) {
            const size_t pf = at(li_, lj_, ix);
            for (int f = 0; f < 6; f++) l[f] += rw(a, RW_C1 + dir * 12 + 6 + f)[pf];
        }
        if (r4_face(dir, i, j, ix, jx)) {                                    // face after: its L
            for (int f = 0; f < 6; f++) l[f] += rw(a, RW_C1 + dir * 12 + f)[k];
        }
    }
    for (int f = 0; f < 6; f++) a.lout[f][k] = l[f];
}

AdjArgs make_adj(const double *const *pn, const double *const *pn1, const double *const *lin, double *const *lout,
                 double *ws, const double *dtp, double *acc, const double *dx, const double *dxm, const double *dy,
                 const double *dym, Metric mx, Metric my, double qav, double gm, double safety, int ix, int jx,
                 int margin) {
    const size_t n = (size_t)ix * jx;
    const size_t na = (n + 63) & ~size_t(63);
    AdjArgs a{};
    for (int f = 0; f < 6; f++) {
        a.p[f] = pn[f];
        a.q[f] = pn1 ? pn1[f] : nullptr;
        a.lin[f] = lin ? const_cast<double *>(lin[f]) : nullptr;
        a.lout[f] = lout[f];
        a.ldu[f] = ws + f * na;
        a.luh[f] = ws + (6 + f) * na;
    }
    a.dtp = dtp;
    a.ldt = acc;
    a.nties = reinterpret_cast<unsigned *>(acc + kAdjStripes);
    a.dx = dx; a.dxm = dxm; a.dy = dy; a.dym = dym;
    a.mx = mx; a.my = my;
    a.qav = qav; a.gm = gm; a.safety = safety;
    a.ix = ix; a.jx = jx; a.margin = margin;
    a.ilo = 2;
    return a;
}

}  // namespace

hipError_t launch_adj2d_step(const double *const *pn, const double *const *pn1, const double *const *lin,
                             double *const *lout, double *ws, const double *dtp, double *acc, const double *pend_dt,
                             const double *pend_acc, double *acc_zero, const double *dx, const double *dxm,
                             const double *dy, const double *dym, Metric mx, Metric my, double qav, double gm,
                             double safety, int ix, int jx, int margin, hipStream_t s) {
    AdjArgs a = make_adj(pn, pn1, lin, lout, ws, dtp, acc, dx, dxm, dy, dym, mx, my, qav, gm, safety, ix, jx, margin);
    a.pend_dt = pend_dt;
    a.pend_acc = pend_acc;
    a.acc_zero = acc_zero;
    const size_t n = (size_t)ix * jx;
    hipLaunchKernelGGL(adj_b1, dim3(xcd_grid(n)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(adj_b2, dim3(xcd_grid((size_t)(ix - 1) * (jx - 1))), dim3(256), 0, s, a);
    hipLaunchKernelGGL(adj_b3, dim3(xcd_grid(n)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_adj2d_dt(const double *const *pn, double *const *lout, double *ws, const double *dtp, double *acc,
                           const double *dx, const double *dy, double gm, double safety, int ix, int jx,
                           hipStream_t s) {
    Metric z{};
    AdjArgs a = make_adj(pn, nullptr, nullptr, lout, ws, dtp, acc, dx, nullptr, dy, nullptr, z, z, 0.0, gm, safety,
                         ix, jx, 1);
    const size_t n = (size_t)ix * jx;
    hipLaunchKernelGGL(adj_dt, dim3((int)((n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

__global__ void key_to_dt(const unsigned long long *key, double *dt) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *dt = dunkey(*key);
}

hipError_t launch_key_to_dt(const unsigned long long *key, double *dt, hipStream_t s) {
    hipLaunchKernelGGL(key_to_dt, dim3(1), dim3(64), 0, s, key, dt);
    return hipGetLastError();
}

__global__ void zero_kernel(double *p, size_t n) {
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        p[k] = 0.0;
}

hipError_t launch_zero(double *p, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(zero_kernel, dim3(64), dim3(256), 0, s, p, n);
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

hipError_t launch_adj2d_roe_step(const double *const *pn, const double *const *pn1, const double *const *ph,
                                 const double *const *lin, double *const *lout, double *ws, const double *dtp,
                                 double *acc, const double *pend_dt, const double *pend_acc, double *acc_zero,
                                 const double *dx, const double *dy, double gm, double safety, int ix, int jx,
                                 int margin, hipStream_t s) {
    const size_t n = (size_t)ix * jx;
    const size_t na = (n + 63) & ~size_t(63);
    RoeAdj r{};
    r.w = ws;
    r.na = na;
    for (int f = 0; f < 6; f++) {
        r.p[f] = pn[f];
        r.lout[f] = lout[f];
    }
    const int slot[6] = {0, 1, 2, 3, 5, 6};   // the Roe engine's slots of ro pr vx vy bx by
    for (int f = 0; f < 6; f++) r.ph[f] = ph ? ph[f] : ws + (RW_PH + slot[f]) * na;
    r.dx = dx; r.dy = dy; r.dtp = dtp; r.ldt = acc; r.acc_zero = acc_zero; r.gm = gm; r.ix = ix; r.jx = jx;
    // the forward's half-step primitives at P^n: recorded, or recomputed
    // (the stage-1 fluxes and the half step) into ws
    RoeArgs f{};
    for (int q = 0; q < 6; q++) f.prim[slot[q]] = pn[q];
    for (int q = 0; q < 8; q++) {
        f.half[q] = ws + (RW_PH + q) * na;
        f.flux[0][q] = ws + (RW_F1 + q) * na;
        f.flux[1][q] = ws + (RW_F1 + 8 + q) * na;
    }
    f.dx = dx; f.dy = dy; f.gm = gm; f.dtp = dtp; f.ix = ix; f.jx = jx; f.kx = 1;
    if (!ph) {
        hipError_t e = launch_roe2d_half(f, s);
        if (e != hipSuccess) return e;
    }
    Metric z{};
    AdjArgs a = make_adj(pn, pn1, lin, lout, ws, dtp, acc, dx, nullptr, dy, nullptr, z, z, 0.0, gm, safety, ix, jx,
                         margin);
    a.ilo = margin + 1;
    a.pend_dt = pend_dt;
    a.pend_acc = pend_acc;
    a.acc_zero = nullptr;
    hipLaunchKernelGGL(adj_b1, dim3(xcd_grid(n)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(roe_r2<0>, dim3(xcd_grid(n)), dim3(256), 0, s, r);
    hipLaunchKernelGGL(roe_r2<1>, dim3(xcd_grid(n)), dim3(256), 0, s, r);
    hipLaunchKernelGGL(roe_r3, dim3(xcd_grid(n)), dim3(256), 0, s, r);
    hipLaunchKernelGGL(roe_r4<0>, dim3(xcd_grid(n)), dim3(256), 0, s, r);
    hipLaunchKernelGGL(roe_r4<1>, dim3(xcd_grid(n)), dim3(256), 0, s, r);
    hipLaunchKernelGGL(roe_r5, dim3(xcd_grid(n)), dim3(256), 0, s, r);
    return hipGetLastError();
}

size_t adj2d_roe_ws_doubles(int ix, int jx) {
    const size_t n = (size_t)ix * jx;
    const size_t na = (n + 63) & ~size_t(63);
    return RW_N * na + 64;
}

size_t adj2d_ws_doubles(int ix, int jx) {
    const size_t n = (size_t)ix * jx;
    const size_t na = (n + 63) & ~size_t(63);
    static_assert(4 + 2 * kAdjAccSlot <= 64, "adjoint accumulator tail");
    return 12 * na + 64;
}

}  // namespace cans
