"""The UTRIP spicule model (cans1d/md_spicule/hdmlw): a 1D flux tube through a
stratified atmosphere with gravity, driven by a random torque at its foot
point, advanced by ``mlw_m_bg`` (cans1d/hdmlw/mlw_m_bg.f).

Host side (numpy, FP64, bit for bit the reference's arithmetic):

* ``model``   -- model.f:1-192: stretched grid, gravity, temperature,
                 hydrostatic density, uniform Bx, flux-tube cross-section.
* ``setup``   -- main.f:63-83: model, grdrdy, bnd, scrdy (sc and rr),
                 bndsc, chkdav; everything the time loop reads.
* ``torque_table``, ``draw_powers`` -- the time-independent parts of
                 trqalfven.f: the torque profile and the Park-Miller stream
                 (common/rangen.f) that replaces RANDOM_NUMBER.

Device side (``SpiculeRuns``): any number of independent realizations (own
seed, optionally own geometry) each running the whole main.f time loop in
one persistent kernel (cans1d_spicule_run).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import Dict, Optional, Sequence

import numpy as np

from .problems import grdrdy

# main.f:168-199 constants
SAFETY = 0.8
DTMIN = 1.0e-10
QAV = 1.0
FLOOR = 1.0e-9
TEND = 100.0
DTOUT = 0.5
NSTOP = 100000

# Park-Miller minimal standard (common/rangen.f:11-15)
PM_A = 16807
PM_M = 2147483647
PM_RNML = 4.656612875245796924105750827168e-10

R4_0075 = float(np.float32(0.075))   # REAL*4 literal 0.075 of trqalfven.f:25

# order of the geometry arrays handed to the device loop
GEOM = ("bx", "bxm", "gx", "gxm", "sc", "dsc", "scm", "dscm", "rr", "rrm", "drr", "drrm", "dx", "dxm")
FIELDS = ("ro", "pr", "vx", "vy", "by")


def _bd_low(q, kind, margin):
    """bdsppx/bdspnx/bdsmpx/bdsmnx/bdfrex with mbnd=0 (cans1d/bc/*.f); 1-based
    logic on a 0-based array."""
    ibnd = 1 + margin
    for i in range(1, margin + 1):
        if kind == "spp":
            q[ibnd - i - 1] = q[ibnd - 1 + i - 1]
        elif kind == "spn":
            q[ibnd - i - 1] = -q[ibnd - 1 + i - 1]
        elif kind == "smp":
            q[ibnd - i - 1] = q[ibnd + i - 1]
        elif kind == "smn":
            q[ibnd - i - 1] = -q[ibnd + i - 1]
        elif kind == "fre":
            q[ibnd - i - 1] = q[ibnd - 1]
    if kind == "smn":
        q[ibnd - 1] = 0.0


def _bd_high(q, kind, margin):
    ix = q.shape[0]
    ibnd = ix - margin
    for i in range(1, margin + 1):
        if kind == "spp":
            q[ibnd + i - 1] = q[ibnd + 1 - i - 1]
        elif kind == "spn":
            q[ibnd + i - 1] = -q[ibnd + 1 - i - 1]
        elif kind == "smp":
            q[ibnd + i - 1] = q[ibnd - i - 1]
        elif kind == "smn":
            q[ibnd + i - 1] = -q[ibnd - i - 1]
        elif kind == "fre":
            q[ibnd + i - 1] = q[ibnd - 1]
    if kind == "smn":
        q[ibnd - 1] = 0.0


def model(ix: int = 1026, margin: int = 1, onebeta: float = 1.0) -> Dict[str, np.ndarray]:
    """md_spicule/hdmlw/model.f:1-192.  Returns x, ro, pr, vx, vy, by, bx,
    bxm, gx, gxm, rr, rrm, sc, scm, dv and gm.  ``onebeta`` is model.f:149's
    constant (1 in the shipped run)."""
    ztr = 15.0
    gm = 5.0 / 3.0
    # grid (model.f:31-66): uniform dx0 up to i1, then stretched by 2 % per cell
    dx0 = 0.04
    i1 = int(math.trunc(2.0 * ztr / dx0))
    apx1 = 1.02
    dxmax1 = 50.0 * dx0
    dxm = [dx0] * (ix + 1)                     # 1-based
    if i1 < ix:
        for i in range(i1 + 1, ix + 1):
            dxm[i] = apx1 * dxm[i - 1]
            if dxm[i] > dxmax1:
                dxm[i] = dxmax1
    for i in range(1, margin):                 # bdsmpx(0, margin-1, dxm)
        dxm[margin - i] = dxm[margin + i]
    izero = margin + 1
    x = [0.0] * (ix + 1)
    x[izero] = dxm[izero] / 2.0
    for i in range(izero + 1, ix + 1):
        x[i] = x[i - 1] + dxm[i - 1]
    for i in range(izero - 1, 0, -1):
        x[i] = x[i + 1] - dxm[i]

    pi = math.acos(-1.0)
    gl0 = 1.0 / gm
    ro0 = 1.0
    pr0 = 1.0 / gm
    te0 = 1.0
    tpho = te0
    wtr = 0.5
    # gravity (model.f:88-101), constant up to x = 400
    zg0, wg0 = 400.0, 10.0
    gx = [0.0] + [-(gl0 * (1.0 - math.tanh((x[i] - zg0) / wg0)) / 2.0) for i in range(1, ix + 1)]
    gxm = list(gx)
    ibnd = 1 + margin                          # bdspnx(0, margin, gx)
    for i in range(1, margin + 1):
        gx[ibnd - i] = -gx[ibnd - 1 + i]
    ibnd = margin                              # bdsmnx(0, margin-1, gxm)
    for i in range(1, margin):
        gxm[ibnd - i] = -gxm[ibnd + i]
    gxm[ibnd] = 0.0
    rr = [0.0] + [1.0] * ix
    rrm = [0.0] + [1.0] * ix
    # temperature and hydrostatic density (model.f:113-144)
    ac = 300.0
    te1d = [0.0] + [tpho * (1.0 + 0.5 * (ac - 1.0) * (math.tanh((x[i] - ztr) / wtr) + 1.0))
                    for i in range(1, ix + 1)]
    ro1d = [0.0] * (ix + 1)
    i = izero + 1
    h = 0.5 * gm * gxm[i - 1] * dxm[i - 1] / 2.0
    ro1d[i] = (te1d[i - 1] + h) / (te1d[i] - h) * ro0
    for i in range(izero + 2, ix + 1):
        h = 0.5 * gm * gxm[i - 1] * dxm[i - 1]
        ro1d[i] = (te1d[i - 1] + h) / (te1d[i] - h) * ro1d[i - 1]
    i = izero
    h = 0.5 * gm * gxm[i] * dxm[i] / 2.0
    ro1d[i] = (te1d[i + 1] - h) / (te1d[i] + h) * ro0
    for i in range(izero - 1, 0, -1):
        h = 0.5 * gm * gxm[i] * dxm[i]
        ro1d[i] = ro1d[i + 1] * (te1d[i + 1] - h) / (te1d[i] + h)
    pr1d = [0.0] + [pr0 * (ro1d[i] / ro0) * (te1d[i] / te0) for i in range(1, ix + 1)]
    # initial state and field (model.f:150-176)
    bx1 = math.sqrt(onebeta * 8.0 * pi * pr1d[1])
    bx = [0.0, bx1] + [bx1 * ((rr[1] / rr[i]) * (rr[1] / rr[i])) for i in range(2, ix + 1)]
    bxm = [0.0] + [0.5 * (bx[i] + bx[i + 1]) for i in range(1, ix)] + [0.0]
    bxm[ix] = bxm[ix - 1]
    sc = [0.0] + [1.0 / bx[i] for i in range(1, ix + 1)]
    scm = [0.0] + [1.0 / bxm[i] for i in range(1, ix + 1)]

    def a(v):
        return np.array(v[1:], dtype=np.float64)

    out = dict(x=a(x), ro=a(ro1d), pr=a(pr1d), vx=np.zeros(ix), vy=np.zeros(ix), by=np.zeros(ix), bx=a(bx),
               bxm=a(bxm), gx=a(gx), gxm=a(gxm), rr=a(rr), rrm=a(rrm), sc=a(sc), scm=a(scm))
    out["dv"] = out["sc"].copy()
    out["gm"] = gm
    return out


def scrdy(dsc, dscm, sc, scm, dx, dxm) -> None:
    """common/scrdy.f:28-34, in place (only elements 2..ix-1 / 2..ix-2)."""
    ix = sc.shape[0]
    dsc[1: ix - 1] = (scm[1: ix - 1] - scm[0: ix - 2]) / dx[1: ix - 1]
    dscm[1: ix - 2] = (sc[2: ix - 1] - sc[1: ix - 2]) / dxm[1: ix - 2]


def bnd(margin, ro, pr, vx, vy, by) -> None:
    """md_spicule/hdmlw/bnd.f:10-20: mirror about the foot-point face (vx
    odd), free at the top."""
    for q, k in ((pr, "spp"), (ro, "spp"), (vx, "spn"), (vy, "spp"), (by, "spp")):
        _bd_low(q, k, margin)
    for q in (pr, ro, vx, vy, by):
        _bd_high(q, "fre", margin)


def bndsc(margin, sc, dsc, scm, dscm) -> None:
    """md_spicule/hdmlw/bndsc.f:10-18."""
    _bd_low(sc, "spp", margin)
    _bd_low(dsc, "spn", margin)
    _bd_low(scm, "smp", margin - 1)
    _bd_low(dscm, "smn", margin - 1)
    _bd_high(sc, "spp", margin)
    _bd_high(dsc, "spn", margin)
    _bd_high(scm, "smp", margin)
    _bd_high(dscm, "smn", margin)


def chkdav(da, vx, floor=FLOOR) -> int:
    """common/chkdav.f:31-39, in place; returns the number of floored cells."""
    m = da < floor
    da[m] = floor
    vx[m] = 0.0
    return int(m.sum())


@dataclass
class Spicule:
    """Everything main.f's time loop reads: state, geometry, constants."""
    ix: int
    margin: int
    gm: float
    x: np.ndarray
    state: Dict[str, np.ndarray]
    geom: Dict[str, np.ndarray]


def setup(ix: int = 1026, margin: int = 1, onebeta: float = 1.0) -> Spicule:
    """main.f:63-83: model, grdrdy, bnd, scrdy(sc), scrdy(rr), bndsc, chkdav.
    The main program's arrays live in static storage, so the elements that
    scrdy/bndsc never write (drr, drrm at the ends) are zero."""
    m = model(ix, margin, onebeta)
    dx, xm, dxm = grdrdy(m["x"])
    st = {k: m[k].copy() for k in FIELDS}
    bnd(margin, st["ro"], st["pr"], st["vx"], st["vy"], st["by"])
    dsc, dscm, drr, drrm = (np.zeros(ix) for _ in range(4))
    scrdy(dsc, dscm, m["sc"], m["scm"], dx, dxm)
    scrdy(drr, drrm, m["rr"], m["rrm"], dx, dxm)
    sc, scm = m["sc"].copy(), m["scm"].copy()
    bndsc(margin, sc, dsc, scm, dscm)
    chkdav(st["ro"], st["vx"])
    chkdav(st["pr"], st["vx"])
    geom = dict(bx=m["bx"], bxm=m["bxm"], gx=m["gx"], gxm=m["gxm"], sc=sc, dsc=dsc, scm=scm, dscm=dscm,
                rr=m["rr"], rrm=m["rrm"], drr=drr, drrm=drrm, dx=dx, dxm=dxm)
    return Spicule(ix=ix, margin=margin, gm=m["gm"], x=m["x"], state=st, geom=geom)


def torque_table(x: np.ndarray) -> np.ndarray:
    """trqalfven.f:23-28: trq(i) = tanh((x(i)-0.75)/0.075)-1, trq(ix) =
    trq(ix-2).  Time-independent, so evaluated once on the host (libm tanh)."""
    ix = x.shape[0]
    t = np.array([math.tanh((float(x[i]) - 0.75) / R4_0075) - 1.0 for i in range(ix)])
    t[ix - 1] = t[ix - 3]
    return t


def draw_powers(ix: int) -> np.ndarray:
    """16807^(i+1) mod (2^31-1) for i < ix: cell i+1 of a trqalfven call
    draws seed*16807^(i+1) mod (2^31-1) from the call's starting seed (one
    draw per cell in cell order); element ix-1 is also the seed's advance
    per call."""
    p = np.empty(ix, dtype=np.int64)
    a = 1
    for i in range(ix):
        a = a * PM_A % PM_M
        p[i] = a
    return p


def rangen(seed: int):
    """common/rangen.f:11-21 (Schrage), returning (value, new seed)."""
    k = seed // 127773
    s = PM_A * (seed - k * 127773) - 2836 * k
    if s < 0:
        s += PM_M
    return PM_RNML * s, s


class SpiculeRuns:
    """``nrun`` independent realizations of the spicule time loop
    (md_spicule/hdmlw/main.f:169-227) on the GPU, each with its own
    Park-Miller seed (and, with a list of set-ups, its own geometry), in one
    persistent kernel (cans1d_spicule_run).

    ``run(nstop, tend)`` continues every realization to ns >= nstop or
    t >= tend; ``snapshots`` holds the states main.f would write to its DAC
    files (every time int(t/dtout) changes)."""

    def __init__(self, sp, seeds: Sequence[int], device="cuda:0", nsnap_max: int = 0, dtout: float = DTOUT,
                 safety: float = SAFETY, dtmin: float = DTMIN, qav: float = QAV, floor: float = FLOOR):
        import torch
        from . import cans
        self._torch, self._cans = torch, cans
        sps = list(sp) if isinstance(sp, (list, tuple)) else [sp]
        self.nrun = len(seeds)
        if len(sps) not in (1, self.nrun):
            raise ValueError("give one set-up, or one per realization")
        for s in seeds:
            if not 1 <= int(s) <= PM_M - 1:
                raise ValueError(f"seed {s} not in 1..2^31-2")
        self.per_run = len(sps) > 1
        s0 = sps[0]
        self.ix, self.margin, self.gm = s0.ix, s0.margin, s0.gm
        if any(s.ix != self.ix or s.margin != self.margin or s.gm != self.gm for s in sps):
            raise ValueError("realizations must share ix, margin and gm")
        self.safety, self.dtmin, self.qav, self.floor, self.dtout = safety, dtmin, qav, floor, dtout
        dev = torch.device(device)
        self.dev = dev
        ix, n = self.ix, self.nrun

        def tile(arrs):
            a = np.concatenate(arrs) if len(arrs) > 1 else arrs[0]
            return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)

        self.f = [tile([s.state[k] for s in sps] if self.per_run else [np.tile(s0.state[k], n)]) for k in FIELDS]
        self.geo = [tile([s.geom[k] for s in sps]) for k in GEOM]
        self.trq = tile([torque_table(s.x) for s in sps])
        self.x = s0.x
        self.t = torch.zeros(n, dtype=torch.float64, device=dev)
        self.ns = torch.zeros(n, dtype=torch.int32, device=dev)
        self.merr = torch.zeros(n, dtype=torch.int32, device=dev)
        self.seed = torch.tensor([int(s) for s in seeds], dtype=torch.int32, device=dev)
        self.nsnap_max = int(nsnap_max)
        if self.nsnap_max:
            self.snap = torch.zeros(n * self.nsnap_max * 5 * ix, dtype=torch.float64, device=dev)
            self.snap_t = torch.zeros(n * self.nsnap_max, dtype=torch.float64, device=dev)
            self.nsnap = torch.zeros(n, dtype=torch.int32, device=dev)
        else:
            self.snap = self.snap_t = self.nsnap = None
        L = cans.lib()
        L.cans1d_spicule_ws_bytes.restype = C.c_size_t
        nb = L.cans1d_spicule_ws_bytes(C.c_int(ix), C.c_int(n), C.c_int(1 if self.per_run else 0))
        self.ws = torch.empty((nb + 7) // 8, dtype=torch.float64, device=dev)
        self.block_threads = 256

    def run(self, nstop: int = NSTOP, tend: float = TEND, block_threads: Optional[int] = None, blocks: int = 0):
        cans = self._cans
        ix, n = self.ix, self.nrun
        P = cans._ptr
        st = (C.c_void_p * 5)(*[P(t, (n * ix,)).value for t in self.f])
        ng = (n if self.per_run else 1) * ix
        geo = (C.c_void_p * len(GEOM))(*[P(t, (ng,)).value for t in self.geo])
        ip = cans._iptr
        snap = P(self.snap) if self.snap is not None else C.c_void_p(0)
        snap_t = P(self.snap_t) if self.snap_t is not None else C.c_void_p(0)
        nsnap = ip(self.nsnap, self._torch.int32, n, "nsnap") if self.nsnap is not None else C.c_void_p(0)
        bt = int(block_threads or self.block_threads)
        cans._check(cans.lib().cans1d_spicule_run(
            C.c_int(n), C.c_int(ix), C.c_int(self.margin), st, geo, C.c_int(1 if self.per_run else 0),
            P(self.trq, (ng,)), cans._d(self.gm), cans._d(self.qav), cans._d(self.safety), cans._d(self.dtmin),
            cans._d(self.floor), cans._d(tend), C.c_int(int(nstop)), P(self.t, (n,)), ip(self.ns, self._torch.int32, n, "ns"),
            ip(self.merr, self._torch.int32, n, "merr"), ip(self.seed, self._torch.int32, n, "seed"),
            cans._d(self.dtout), snap, snap_t, C.c_int(self.nsnap_max), nsnap, C.c_int(bt), C.c_int(int(blocks)),
            P(self.ws), cans._stream()))

    def state(self, r: int) -> Dict[str, np.ndarray]:
        s = slice(r * self.ix, (r + 1) * self.ix)
        return {k: t[s].cpu().numpy() for k, t in zip(FIELDS, self.f)}

    def snapshots(self, r: int):
        """(times, states[k, 5, ix]) written for realization r by the last run."""
        if self.nsnap is None:
            return np.zeros(0), np.zeros((0, 5, self.ix))
        k = int(self.nsnap[r].item())
        t = self.snap_t[r * self.nsnap_max: r * self.nsnap_max + k].cpu().numpy()
        s = self.snap.view(self.nrun, self.nsnap_max, 5, self.ix)[r, :k].cpu().numpy()
        return t, s

    def info(self, r: int):
        return dict(t=float(self.t[r].item()), ns=int(self.ns[r].item()), merr=int(self.merr[r].item()),
                    seed=int(self.seed[r].item()))


def mlw_m_bg(ro, pr, vx, vy, by, geom, dt, qav, gm):
    """One mlw_m_bg call on device tensors (cans1d_mlw_m_bg); ``geom``: dict
    of the 14 geometry tensors (GEOM)."""
    from . import cans
    ix = ro.numel()
    P = lambda t, k="array": cans._ptr(t, (ix,), k)  # noqa: E731
    cans._check(cans.lib().cans1d_mlw_m_bg(
        P(ro), P(pr), P(vx), P(vy), P(by), P(geom["bx"]), P(geom["bxm"]), cans._d(dt), cans._d(qav), cans._d(gm),
        *[P(geom[k], k) for k in GEOM[2:]], C.c_int(ix), cans._stream()))


def trqalfven(vy, dt, trq, rr, seed: int) -> int:
    """One trqalfven call on device tensors (cans1d_trqalfven); returns the
    advanced seed."""
    from . import cans
    ix = vy.numel()
    s = C.c_int(int(seed))
    cans._check(cans.lib().cans1d_trqalfven(cans._ptr(vy, (ix,)), cans._d(dt), cans._ptr(trq, (ix,)),
                                            cans._ptr(rr, (ix,)), C.c_int(ix), C.byref(s), cans._stream()))
    return s.value


def chkdav_dev(da, vx, floor=FLOOR) -> int:
    """chkdav on device tensors (cans1d_chkdav); returns n_events."""
    from . import cans
    ix = da.numel()
    n = C.c_int()
    cans._check(cans.lib().cans1d_chkdav(C.byref(n), cans._ptr(da, (ix,)), cans._ptr(vx, (ix,)), cans._d(floor),
                                         C.c_int(ix), cans._stream()))
    return n.value
