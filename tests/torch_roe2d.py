"""One step of the periodic 2D Roe-MUSCL loop (cfl_m -> roe_m -> bdperx,
bdpery) as vectorised torch float64 operations, following
cans2d/hdroe/roe_m.f, tvdminmod.f and cans2d/common/cfl_m.f.  Test
infrastructure: autograd of this step is the reference vector-Jacobian
product the hand-written Roe adjoint is checked against.

The face flux is the CANS kernels' roeflux12 (utrip_amd/csrc/cans_roeflux.hpp,
a statement-for-statement restatement of cans2d/hdroe/roeflux_m.f:48-242):
its statements are read from that file and evaluated with torch operations,
so that autograd differentiates the same expressions, branch for branch
(torch.where for MAX, copysign for SIGN).  Tensors are (jx, ix) with ghost
width m >= 3."""
import math
import re
from pathlib import Path

import torch

PI = math.pi
PI4 = 4.0 * PI
PI4I = 1.0 / PI4
PI8I = 5.0e-1 * PI4I
SQH = 0.70710678118654757

_HPP = Path(__file__).resolve().parents[1] / "utrip_amd" / "csrc" / "cans_roeflux.hpp"


def _body(name):
    src = _HPP.read_text()
    i = src.index(f"CANS_DEV void {name}(")
    i = src.index("{", i) + 1
    depth, j = 1, i
    while depth:
        depth += {"{": 1, "}": -1}.get(src[j], 0)
        j += 1
    return src[i:j - 1]


def _translate(body):
    out = []
    for st in re.sub(r"//[^\n]*", "", body).split(";"):
        st = " ".join(st.split())
        if not st or st.startswith("(void)"):
            continue
        st = re.sub(r"^const T ", "", st)
        st = re.sub(r"T\(RoePi<T>::(\w+)\)", lambda m: {"pi4d": "PI4", "pi4id": "PI4I", "pi8id": "PI8I",
                                                       "sqh": "SQH"}[m.group(1)], st)
        st = re.sub(r"\bT\(([^()]*)\)", r"(\1)", st)
        st = re.sub(r"^out\[(\d)\] = ", r"out\1 = ", st)
        out.append(st)
    return out


def _fmax(a, b):
    a = torch.as_tensor(a, dtype=torch.float64)
    b = torch.as_tensor(b, dtype=torch.float64)
    return torch.where(a > b, a, b)


def _sign(a, b):
    b = torch.as_tensor(b, dtype=torch.float64)
    return torch.copysign(torch.full_like(b, abs(float(a))), b)


_ROE_SRC = None


def roeflux12(gm, rhol, prl, vxl, vyl, bxl, byl, rhor, prr, vxr, vyr, bxr, byr):
    """The five fluxes (fro, fee, frx, fry, fby) of one sweep frame."""
    global _ROE_SRC
    if _ROE_SRC is None:
        lines = _translate(_body("roeflux12"))
        # the statements that assign across a comma become one per line
        code = []
        for ln in lines:
            parts, depth, cur = [], 0, ""
            for ch in ln:
                depth += {"(": 1, ")": -1}.get(ch, 0)
                if ch == "," and depth == 0:
                    parts.append(cur)
                    cur = ""
                else:
                    cur += ch
            parts.append(cur)
            code += [p.strip() for p in parts]
        _ROE_SRC = compile("\n".join(code), "roeflux12", "exec")
    env = dict(gm=gm, rhol=rhol, prl=prl, vxl=vxl, vyl=vyl, bxl=bxl, byl=byl, rhor=rhor, prr=prr, vxr=vxr, vyr=vyr,
               bxr=bxr, byr=byr, tsqrt=torch.sqrt, tabs=torch.abs, fmaxF=_fmax, signF=_sign, PI4=PI4, PI4I=PI4I,
               PI8I=PI8I, SQH=SQH)
    exec(_ROE_SRC, env)
    return [env[f"out{k}"] for k in range(5)]


def _flmt(a, b):   # tvdminmod.f statement function
    s = torch.copysign(torch.ones_like(a), a)
    x = b * s
    y = torch.abs(a)
    m = torch.where(x < y, x, y)
    return torch.where(0.0 > m, torch.zeros_like(m), m) * s


def fill(t, m):
    """bdperx then bdpery of ghost width m on a (jx, ix) tensor."""
    ix, jx = t.shape[1], t.shape[0]
    t = torch.cat([t[:, ix - 2 * m: ix - m], t[:, m: ix - m], t[:, m: 2 * m]], dim=1)
    return torch.cat([t[jx - 2 * m: jx - m, :], t[m: jx - m, :], t[m: 2 * m, :]], dim=0)


def _cons(P, gm):
    ro, pr, vx, vy, bx, by = P
    return [ro, pr / (gm - 1.0) + 0.5 * ro * (vx * vx + vy * vy) + PI8I * (bx * bx + by * by), ro * vx, ro * vy, bx, by]


def _prim(U, gm):
    ro = U[0]
    vx = U[2] / ro
    vy = U[3] / ro
    pr = (gm - 1.0) * (U[1] - 0.5 * ro * (vx * vx + vy * vy) - PI8I * (U[4] * U[4] + U[5] * U[5]))
    return [ro, pr, vx, vy, U[4], U[5]]


def _faces(P, dir, L, R, gm):
    """Fluxes of direction dir from left/right state lists, as U slots
    (ro, ee, rx, ry, bx, by; the normal field's slot None)."""
    vn, vt, bn, bt = (2, 3, 4, 5) if dir == 0 else (3, 2, 5, 4)
    o = roeflux12(gm, L[0], L[1], L[vn], L[vt], L[bn], L[bt], R[0], R[1], R[vn], R[vt], R[bn], R[bt])
    F = [None] * 6
    F[0], F[1], F[vn], F[vt], F[bt] = o[0], o[1], o[2], o[3], o[4]
    return F


def step(P, dx, dy, gm, safety, m):
    """P: 6 tensors ro,pr,vx,vy,bx,by (jx, ix), ghosts filled; returns the
    next state (ghosts filled) and dt."""
    ro, pr, vx, vy, bx, by = P
    jx, ix = ro.shape
    S = vx * vx + vy * vy + gm * pr / ro + (bx * bx + by * by) * PI4I / ro
    dmin = torch.minimum(dx[None, :], dy[:, None])
    dt = torch.amin(safety * dmin / torch.sqrt(S))
    U = _cons(P, gm)
    # stage 1: x faces between columns c, c+1 (c = 0..ix-2), y faces between rows
    F1x = _faces(P, 0, [q[:, :-1] for q in P], [q[:, 1:] for q in P], gm)
    F1y = _faces(P, 1, [q[:-1, :] for q in P], [q[1:, :] for q in P], gm)
    # half step on cells [2, n-1]^2: x sweep, then y sweep
    ci = (slice(1, -1), slice(1, -1))
    Uh = [u[ci] for u in U]
    dxi = dx[1:-1][None, :]
    dyi = dy[1:-1][:, None]
    for q in (0, 1, 2, 3, 5):
        Uh[q] = Uh[q] + 0.5 * dt * ((F1x[q][1:-1, :-1] - F1x[q][1:-1, 1:]) / dxi)
    for q in (0, 1, 2, 3, 4):
        Uh[q] = Uh[q] + 0.5 * dt * ((F1y[q][:-1, 1:-1] - F1y[q][1:, 1:-1]) / dyi)
    ph_in = _prim(Uh, gm)
    Ph = []
    for q in range(6):
        t = P[q]
        mid = torch.cat([t[1:-1, :1], ph_in[q], t[1:-1, -1:]], dim=1)
        Ph.append(torch.cat([t[:1, :], mid, t[-1:, :]], dim=0))
    # stage 2: minmod states; x faces between columns c, c+1 for c = 1..ix-3
    Lx, Rx, Ly, Ry = [], [], [], []
    for q in range(6):
        d = Ph[q]
        d0, d1, d2, d3 = d[:, 0:-3], d[:, 1:-2], d[:, 2:-1], d[:, 3:]
        Lx.append(d1 + 0.5 * _flmt(d2 - d1, d1 - d0))
        Rx.append(d2 - 0.5 * _flmt(d2 - d1, d3 - d2))
        d0, d1, d2, d3 = d[0:-3, :], d[1:-2, :], d[2:-1, :], d[3:, :]
        Ly.append(d1 + 0.5 * _flmt(d2 - d1, d1 - d0))
        Ry.append(d2 - 0.5 * _flmt(d2 - d1, d3 - d2))
    F2x = _faces(Ph, 0, Lx, Rx, gm)    # (jx, ix-3): face c+1 at column c
    F2y = _faces(Ph, 1, Ly, Ry, gm)    # (jx-3, ix)
    # update on cells [3, n-2]^2
    cu = (slice(2, -2), slice(2, -2))
    Un = [u[cu] for u in U]
    dxu = dx[2:-2][None, :]
    dyu = dy[2:-2][:, None]
    for q in (0, 1, 2, 3, 5):
        Un[q] = Un[q] + dt * ((F2x[q][2:-2, :-1] - F2x[q][2:-2, 1:]) / dxu)
    for q in (0, 1, 2, 3, 4):
        Un[q] = Un[q] + dt * ((F2y[q][:-1, 2:-2] - F2y[q][1:, 2:-2]) / dyu)
    pu = _prim(Un, gm)
    out = []
    for q in range(6):
        t = P[q]
        mid = torch.cat([t[2:-2, :2], pu[q], t[2:-2, -2:]], dim=1)
        out.append(fill(torch.cat([t[:2, :], mid, t[-2:, :]], dim=0), m))
    return out, dt
