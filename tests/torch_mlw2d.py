"""One step of the periodic 2D MLW loop (cfl_m -> mlw_m -> bdperx, bdpery) as
plain vectorised torch float64 operations, following cans2d/hdmlw/mlw_m.f,
mlwhalf.f, mlwfull.f, mlwartv.f and cans2d/common/cfl_m.f term by term.  Test
infrastructure: autograd of this step is the reference vector-Jacobian
product the hand-written adjoint step is checked against (the values agree
with the CANS kernels to rounding, not bit for bit).  Tensors are (jx, ix),
ghost width 1."""
import math

import torch

PI = math.pi
PI8I = 1.0 / PI / 8.0
PI4I = 1.0 / PI / 4.0
EPS_AV = float(torch.tensor(1.0e-4, dtype=torch.float32).double())   # REAL*4 1.0e-4


def _cell(ro, pr, vx, vy, bx, by, gm):
    vv = vx * vx + vy * vy
    bb = bx * bx + by * by
    ez = -vx * by + vy * bx
    ep = pr * gm / (gm - 1.0) + 0.5 * ro * vv
    z = torch.zeros_like(ro)
    u = [ro, pr / (gm - 1.0) + 0.5 * ro * vv + PI8I * bb, ro * vx, ro * vy, bx, by]
    f = [ro * vx, ep * vx - by * ez * PI4I, ro * vx * vx + pr + PI8I * (by * by - bx * bx),
         ro * vy * vx - PI4I * by * bx, z, -ez]
    g = [ro * vy, ep * vy + bx * ez * PI4I, ro * vx * vy - PI4I * bx * by,
         ro * vy * vy + pr + PI8I * (bx * bx - by * by), ez, z]
    return u, f, g


def _prims(u, gm):
    ro = u[0]
    vx = u[2] / ro
    vy = u[3] / ro
    pr = (gm - 1.0) * (u[1] - 0.5 * ro * (vx * vx + vy * vy) - PI8I * (u[4] * u[4] + u[5] * u[5]))
    return [ro, pr, vx, vy, u[4], u[5]]


def _fill(t):
    t = torch.cat([t[:, -2:-1], t[:, 1:-1], t[:, 1:2]], dim=1)
    return torch.cat([t[-2:-1, :], t[1:-1, :], t[1:2, :]], dim=0)


def step(P, dx, dxm, dy, dym, gm, qav, safety):
    """P: 6 tensors ro,pr,vx,vy,bx,by (jx, ix) with ghosts; returns the next
    state (ghosts filled) and dt."""
    ro, pr, vx, vy, bx, by = P
    # cfl_m over every cell (cans2d/common/cfl_m.f:46-55); amin spreads ties evenly
    S = vx * vx + vy * vy + gm * pr / ro + (bx * bx + by * by) * PI4I / ro
    dmin = torch.minimum(dx[None, :], dy[:, None])
    dt = torch.amin(safety * dmin / torch.sqrt(S))
    u, f, g = _cell(ro, pr, vx, vy, bx, by, gm)
    # corners (mlwhalf.f:45-52): A=(j,i) B=(j,i+1) C=(j+1,i) D=(j+1,i+1)
    dxim_c = (1.0 / dxm[:-1])[None, :]
    dyim_c = (1.0 / dym[:-1])[:, None]
    uh = []
    for q in range(6):
        A, B, C, D = u[q][:-1, :-1], u[q][:-1, 1:], u[q][1:, :-1], u[q][1:, 1:]
        fA, fB, fC, fD = f[q][:-1, :-1], f[q][:-1, 1:], f[q][1:, :-1], f[q][1:, 1:]
        gA, gB, gC, gD = g[q][:-1, :-1], g[q][:-1, 1:], g[q][1:, :-1], g[q][1:, 1:]
        uh.append(0.25 * (B + A + D + C) - dt * (dxim_c * (fB - fA + fD - fC) / 2.0)
                  - dt * (dyim_c * (gC - gA + gD - gB) / 2.0))
    _, H, K = _cell(*_prims(uh, gm), gm)
    # update of the interior (mlwhalf centre term, mlwfull, mlwartv)
    dxi = (1.0 / dx[1:-1])[None, :]
    dyi = (1.0 / dy[1:-1])[:, None]
    dxim = (1.0 / dxm[1:-1])[None, :]
    dxim1 = (1.0 / dxm[:-2])[None, :]
    dyim = (1.0 / dym[1:-1])[:, None]
    dyim1 = (1.0 / dym[:-2])[:, None]
    ux0 = (0.5 * dxm[1:-1] / dx[1:-1])[None, :]
    ux1 = (0.5 * dxm[:-2] / dx[1:-1])[None, :]
    uy0 = (0.5 * dym[1:-1] / dy[1:-1])[:, None]
    uy1 = (0.5 * dym[:-2] / dy[1:-1])[:, None]
    c = (slice(1, -1), slice(1, -1))
    E, W, N, Sx = (slice(1, -1), slice(2, None)), (slice(1, -1), slice(None, -2)), \
        (slice(2, None), slice(1, -1)), (slice(None, -2), slice(1, -1))
    qxc = qav * dxm[1:-1][None, :] * torch.clamp(torch.abs(vx[E] - vx[c]) - EPS_AV, min=0.0)
    qxw = qav * dxm[:-2][None, :] * torch.clamp(torch.abs(vx[c] - vx[W]) - EPS_AV, min=0.0)
    qyc = qav * dym[1:-1][:, None] * torch.clamp(torch.abs(vy[N] - vy[c]) - EPS_AV, min=0.0)
    qys = qav * dym[:-2][:, None] * torch.clamp(torch.abs(vy[c] - vy[Sx]) - EPS_AV, min=0.0)
    hdt = 0.5 * dt
    un = []
    for q in range(6):
        du = -hdt * (0.5 * dxi * (f[q][E] - f[q][W])) - hdt * (0.5 * dyi * (g[q][N] - g[q][Sx]))
        hSW, hSE, hNW, hNE = H[q][:-1, :-1], H[q][:-1, 1:], H[q][1:, :-1], H[q][1:, 1:]
        kSW, kSE, kNW, kNE = K[q][:-1, :-1], K[q][:-1, 1:], K[q][1:, :-1], K[q][1:, 1:]
        du = du - hdt * (dxi * (uy0 * (hSE - hSW) + uy1 * (hNE - hNW))) \
                - hdt * (dyi * (ux0 * (kNW - kSW) + ux1 * (kNE - kSE)))
        du = du + dt * (dxi * (qxc * dxim * (u[q][E] - u[q][c]) - qxw * dxim1 * (u[q][c] - u[q][W]))) \
                + dt * (dyi * (qyc * dyim * (u[q][N] - u[q][c]) - qys * dyim1 * (u[q][c] - u[q][Sx])))
        un.append(u[q][c] + du)
    Pn = _prims(un, gm)
    out = []
    for q in range(6):
        t = P[q].clone()
        t = torch.cat([t[:1, :], torch.cat([t[1:-1, :1], Pn[q], t[1:-1, -1:]], dim=1), t[-1:, :]], dim=0)
        out.append(_fill(t))
    return out, dt
