"""Gradients through the periodic 2D MLW time loop (cans2d/hdmlw/mlw_m.f with
bdperx/bdpery and cfl_m, the loop of cans2d/md_mhdshktb/hdmlw/main.f:150-206)
by a hand-written reverse mode on the GPU (cans2d_mlw_adjoint_step).

``PeriodicRoeAdjoint`` is the same for the 2D Roe-MUSCL loop (every state
kept).  ``PeriodicMlwAdjoint`` runs the forward loop on the resident fused
kernel.
With ``ckpt = 0`` it keeps every state of the trajectory in device memory
(7 fields x (nsteps+1) states: 2.9 GB for 200 steps at 512^2) and its
``backward`` is one reverse sweep; with ``ckpt > 0`` it keeps a checkpoint
every ``ckpt`` steps and the backward recomputes each segment from its
checkpoint (on a second stream) before sweeping it, except the last two
segments, whose states the forward records as it runs them into the two
segment buffers (the backward sweeps them first: the last one's
recomputation would have nothing to overlap with, and the other's would
compete with that sweep for the CUs).
``FinalDensity`` wraps this as a ``torch.autograd.Function`` of the initial
velocities (interior) and a uniform pressure scale ``beta``.
"""
from __future__ import annotations

import ctypes as C
from typing import List

import torch

from . import cans
from .cans import _check, _d, _ptr, _ptr_array, _stream, lib


class PeriodicMlwAdjoint:
    """Forward/backward driver for ``nsteps`` steps of the periodic 2D MLW
    loop on a grid with ghost width ``margin`` (state tensors (jx, ix))."""

    def __init__(self, dx, dxm, dy, dym, qav, gm, safety, margin, nsteps, ckpt=20, dtmin=1e-10):
        if ckpt < 0:
            raise ValueError("ckpt must be >= 0 (0: keep every state)")
        self.grid = (dx, dxm, dy, dym)
        self.ix, self.jx = dx.numel(), dy.numel()
        self.qav, self.gm, self.safety, self.margin = qav, gm, safety, margin
        self.nsteps, self.ckpt, self.dtmin = nsteps, ckpt, dtmin
        dev = dx.device
        self.metric = torch.empty(lib().cans2d_metric_ws_size(self.ix, self.jx), dtype=torch.float64, device=dev)
        _check(lib().cans2d_metric_fill(_ptr(dx), _ptr(dxm), C.c_int(self.ix), _ptr(dy), _ptr(dym), C.c_int(self.jx),
                                        _ptr(self.metric), _stream()))
        lib().cans2d_adj_ws_bytes.restype = C.c_size_t
        nb = lib().cans2d_adj_ws_bytes(C.c_int(self.ix), C.c_int(self.jx))
        self.ws = torch.empty((nb + 7) // 8, dtype=torch.float64, device=dev)
        self.ckpts: List[List[torch.Tensor]] = []

    def _run(self, state):
        dx, dxm, dy, dym = self.grid
        run = cans.MlwRun2D([t.clone() for t in state], dx, dxm, dy, dym, self.qav, self.gm, self.safety,
                            margin=self.margin)
        run.reset(dtmin=self.dtmin)
        return run

    @staticmethod
    def _cur(run, k):
        return run.a if k % 2 == 0 else run.b

    def _forward_all(self, state):
        """Every state of the trajectory into self._all (cans2d_mlw_record);
        the dt of each step into self._dt_all."""
        dx, dxm, dy, dym = self.grid
        dev = self.ws.device
        if getattr(self, "_all", None) is None:
            self._all = [[torch.empty_like(t) for t in state] for _ in range(self.nsteps + 1)]
            self._dt_all = torch.empty(max(self.nsteps, 1), dtype=torch.float64, device=dev)
            self._clock_all = cans.Clock()
            self._mws_all = torch.empty_like(self.metric)
            self._all_p = _ptr_array([t for k in range(self.nsteps + 1) for t in self._all[k]],
                                     (self.jx, self.ix))
        for dst, src in zip(self._all[0], state):
            dst.copy_(src)
        self._clock_all.reset(dtmin=self.dtmin)
        _check(lib().cans2d_mlw_record(self._all_p, C.c_int(self.nsteps), self._clock_all.handle, _d(self.qav),
                                       _d(self.gm), _d(self.safety), _ptr(dx), _ptr(dxm), C.c_int(self.ix), _ptr(dy),
                                       _ptr(dym), C.c_int(self.jx), C.c_int(self.margin), _ptr(self._mws_all),
                                       _ptr(self._dt_all), _stream()))
        self._clock_all.issued = self.nsteps
        info = self._clock_all.read()
        if info["merr"] != 0 or info["nexec"] != self.nsteps:
            raise cans.CansError(f"forward run stopped early: {info}")
        return [t.clone() for t in self._all[self.nsteps]]

    def _segment_buffers(self, like):
        """Two segment buffers of ckpt+1 states each, their clocks, dt logs
        and metric workspaces, and the second stream (allocated once)."""
        if getattr(self, "_seg", None) is None:
            dev = self.ws.device
            self._seg = [[[torch.empty_like(t) for t in like] for _ in range(self.ckpt + 1)] for _ in range(2)]
            self._clocks = [cans.Clock(), cans.Clock()]
            self._dts = [torch.empty(self.ckpt, dtype=torch.float64, device=dev) for _ in range(2)]
            self._mws = [torch.empty_like(self.metric) for _ in range(2)]
            self._rec_stream = torch.cuda.Stream(device=dev)
            self._seg_p = [None, None]
            self._last_recorded = False

    def _record_segment(self, c, b):
        """States of segment c (from its checkpoint, min(ckpt, steps left)
        steps) and their dt into segment buffer b, on the current stream
        (cans2d_mlw_record).  Returns the step count and the state pointers."""
        dx, dxm, dy, dym = self.grid
        n = min(self.ckpt, self.nsteps - c * self.ckpt)
        for dst, src in zip(self._seg[b][0], self.ckpts[c]):
            dst.copy_(src)
        states = _ptr_array([t for k in range(n + 1) for t in self._seg[b][k]], (self.jx, self.ix))
        self._seg_p[b] = states
        self._clocks[b].reset(dtmin=self.dtmin)
        _check(lib().cans2d_mlw_record(states, C.c_int(n), self._clocks[b].handle, _d(self.qav), _d(self.gm),
                                       _d(self.safety), _ptr(dx), _ptr(dxm), C.c_int(self.ix), _ptr(dy),
                                       _ptr(dym), C.c_int(self.jx), C.c_int(self.margin),
                                       _ptr(self._mws[b]), _ptr(self._dts[b]), _stream()))
        return n, states

    def forward(self, state):
        """Run nsteps from ``state`` (7 tensors ro,pr,vx,vy,bx,by,az, ghosts
        filled); keep every state (ckpt = 0) or the checkpoints; return the
        final state (clones).  With checkpoints, the segments up to the last
        two run on the resident fused kernel and the last two are recorded
        into the segment buffers (the last into buffer 0), by the same
        kernels as the backward's recomputation, so the same states bit for
        bit."""
        if self.ckpt == 0:
            return self._forward_all(state)
        self.ckpts = []
        run = self._run(state)
        self._segment_buffers(state)
        nseg = max(1, (self.nsteps + self.ckpt - 1) // self.ckpt)
        npre = min(2, nseg)
        done = 0
        for _ in range(nseg - npre):
            self.ckpts.append([t.clone() for t in self._cur(run, done)])
            k = min(self.ckpt, self.nsteps - done)
            run.steps(k)
            done += k
        if done > 0:
            info = run.clock.read()
            if info["merr"] != 0 or info["nexec"] != done:
                raise cans.CansError(f"forward run stopped early: {info}")
        self.ckpts.append([t.clone() for t in self._cur(run, done)])
        self._last_recorded = False
        self._pre = {}
        for c in range(nseg - npre, nseg):
            b = 0 if c == nseg - 1 else 1
            n, _ = self._record_segment(c, b)
            self._clocks[b].issued = n
            info = self._clocks[b].read()
            if info["merr"] != 0 or info["nexec"] != n:
                raise cans.CansError(f"forward run stopped early (segment {c}): {info}")
            self._pre[c] = b
            if c + 1 < nseg:
                self.ckpts.append([t.clone() for t in self._seg[b][n]])
        self._last_ev = torch.cuda.Event()
        self._last_ev.record(torch.cuda.current_stream(self.ws.device))
        self._last_recorded = True
        return [t.clone() for t in self._seg[0][n]]

    def backward(self, lam):
        """``lam``: dL/dP of the final state on every cell (6 tensors
        ro,pr,vx,vy,bx,by).  Returns dL/dP of the initial state, folded onto
        the interior (the gradient with respect to the interior values the
        ghosts are copies of).

        Segments are processed last to first.  Each is recomputed from its
        checkpoint into one of two segment buffers (cans2d_mlw_record, on a
        second stream) while the reverse sweep of the previous segment runs
        on the caller's stream (cans2d_mlw_adjoint_run); events order a
        buffer's record before its sweep and its sweep before its reuse.
        The last two segments' states are those the forward recorded into
        the buffers, the first time the backward follows that forward (a
        second backward finds the buffers reused and recomputes them)."""
        dx, dxm, dy, dym = self.grid
        shape = (self.jx, self.ix)
        lam = [t.clone().contiguous() for t in lam]
        tmp = [torch.empty_like(t) for t in lam]
        dev = self.ws.device
        if self.ckpt == 0:   # every state kept: one reverse sweep
            lam_p, tmp_p = _ptr_array(lam, shape), _ptr_array(tmp, shape)
            _check(lib().cans2d_mlw_adjoint_run(self._all_p, _ptr(self._dt_all), C.c_int(self.nsteps), lam_p, tmp_p,
                                                _d(self.qav), _d(self.gm), _d(self.safety), _ptr(dx), _ptr(dxm),
                                                C.c_int(self.ix), _ptr(dy), _ptr(dym), C.c_int(self.jx),
                                                C.c_int(self.margin), _ptr(self.metric), _ptr(self.ws), _stream()))
            _check(lib().cans2d_periodic_fold(lam_p, C.c_int(self.ix), C.c_int(self.jx), C.c_int(self.margin),
                                              _stream()))
            return lam
        self._segment_buffers(self.ckpts[0])
        main = torch.cuda.current_stream(dev)
        rec = self._rec_stream
        ev_rec = [torch.cuda.Event(), torch.cuda.Event()]
        ev_sweep = [torch.cuda.Event(), torch.cuda.Event()]
        used = [False, False]
        lam_p, tmp_p = _ptr_array(lam, shape), _ptr_array(tmp, shape)
        order = list(reversed(range(len(self.ckpts))))

        pre = self._pre if self._last_recorded else {}
        self._last_recorded = False

        def record(c, b):
            if pre.get(c) == b:
                # recorded by the forward: the sweep waits on the forward's
                # event (each buffer keeps its own event object, so that a
                # later record into the other buffer cannot move this wait)
                main.wait_event(self._last_ev)
                ev_rec[b].record(main)
                return min(self.ckpt, self.nsteps - c * self.ckpt), self._seg_p[b]
            if used[b]:
                rec.wait_event(ev_sweep[b])
            else:
                rec.wait_stream(main)       # checkpoints and buffers were written on the caller's stream
            with torch.cuda.stream(rec):
                n, states = self._record_segment(c, b)
                ev_rec[b].record(rec)
            return n, states

        pending = record(order[0], 0)
        for idx, c in enumerate(order):
            b = idx % 2
            n, states = pending
            if idx + 1 < len(order):
                pending = record(order[idx + 1], 1 - b)
            main.wait_event(ev_rec[b])
            _check(lib().cans2d_mlw_adjoint_run(states, _ptr(self._dts[b]), C.c_int(n), lam_p, tmp_p, _d(self.qav),
                                                _d(self.gm), _d(self.safety), _ptr(dx), _ptr(dxm), C.c_int(self.ix),
                                                _ptr(dy), _ptr(dym), C.c_int(self.jx), C.c_int(self.margin),
                                                _ptr(self.metric), _ptr(self.ws), _stream()))
            ev_sweep[b].record(main)
            used[b] = True
        _check(lib().cans2d_periodic_fold(lam_p, C.c_int(self.ix), C.c_int(self.jx), C.c_int(self.margin), _stream()))
        rec.wait_stream(main)   # nothing on the record stream outlives the call
        return lam


class PeriodicRoeAdjoint:
    """Forward/backward driver for ``nsteps`` steps of the periodic 2D
    Roe-MUSCL loop (cans2d/hdroe/roe_m.f with cfl_m, bdperx, bdpery; ghost
    width ``margin`` >= 3; state tensors (jx, ix)).  The forward records the
    trajectory on the device (cans2d_roe_record: every state, every step's
    half-step primitives and dt); the backward is one reverse sweep
    (cans2d_roe_adjoint_run).  Same interface as PeriodicMlwAdjoint."""

    def __init__(self, dx, dy, gm, safety, margin, nsteps, dtmin=1e-10):
        if margin < 3:
            raise ValueError("the Roe engine needs margin >= 3")
        self.dx, self.dy = dx, dy
        self.ix, self.jx = dx.numel(), dy.numel()
        self.gm, self.safety, self.margin, self.nsteps, self.dtmin = gm, safety, margin, nsteps, dtmin
        dev = dx.device
        lib().cans2d_roe_adj_ws_bytes.restype = C.c_size_t
        nb = lib().cans2d_roe_adj_ws_bytes(C.c_int(self.ix), C.c_int(self.jx))
        self.ws = torch.empty((nb + 7) // 8, dtype=torch.float64, device=dev)
        lib().cans_run_ws_bytes.restype = C.c_size_t
        nr = lib().cans_run_ws_bytes(C.c_int(2), C.c_int(cans.ENGINE_ROE), C.c_int(self.ix), C.c_int(self.jx),
                                     C.c_int(1))
        self.run_ws = torch.empty((nr + 7) // 8, dtype=torch.float64, device=dev)
        self.clock = cans.Clock()
        self._all = None
        self.dt_log = torch.empty(max(nsteps, 1), dtype=torch.float64, device=dev)

    def forward(self, state):
        """Run nsteps from ``state`` (7 tensors ro,pr,vx,vy,bx,by,az, ghosts
        filled); keep every state and half step; return the final state
        (clones)."""
        shape = (self.jx, self.ix)
        if self._all is None:
            self._all = [[torch.empty_like(t) for t in state] for _ in range(self.nsteps + 1)]
            self._half = [[torch.empty_like(state[0]) for _ in range(6)] for _ in range(self.nsteps)]
            self._all_p = _ptr_array([t for k in range(self.nsteps + 1) for t in self._all[k]], shape)
            self._half_p = _ptr_array([t for k in range(self.nsteps) for t in self._half[k]], shape)
        for dst, src in zip(self._all[0], state):
            dst.copy_(src)
        self.clock.reset(dtmin=self.dtmin)
        _check(lib().cans2d_roe_record(self._all_p, self._half_p, C.c_int(self.nsteps), self.clock.handle,
                                       _d(self.gm), _d(self.safety), _ptr(self.dx), _ptr(self.dy), C.c_int(self.ix),
                                       C.c_int(self.jx), C.c_int(self.margin), _ptr(self.run_ws), _ptr(self.dt_log),
                                       _stream()))
        self.clock.issued = self.nsteps
        info = self.clock.read()
        if info["merr"] != 0 or info["nexec"] != self.nsteps:
            raise cans.CansError(f"forward run stopped early: {info}")
        return [t.clone() for t in self._all[self.nsteps]]

    def backward(self, lam, recompute_half=False):
        """``lam``: dL/dP of the final state on every cell (6 tensors
        ro,pr,vx,vy,bx,by).  Returns dL/dP of the initial state folded onto
        the interior.  recompute_half: recompute each step's half step in
        the sweep instead of reading the recorded one (same result)."""
        shape = (self.jx, self.ix)
        lam = [t.clone().contiguous() for t in lam]
        tmp = [torch.empty_like(t) for t in lam]
        lam_p, tmp_p = _ptr_array(lam, shape), _ptr_array(tmp, shape)
        _check(lib().cans2d_roe_adjoint_run(self._all_p, None if recompute_half else self._half_p, _ptr(self.dt_log),
                                            C.c_int(self.nsteps), lam_p, tmp_p, _d(self.gm), _d(self.safety),
                                            _ptr(self.dx), _ptr(self.dy), C.c_int(self.ix), C.c_int(self.jx),
                                            C.c_int(self.margin), _ptr(self.ws), _stream()))
        _check(lib().cans2d_periodic_fold(lam_p, C.c_int(self.ix), C.c_int(self.jx), C.c_int(self.margin), _stream()))
        return lam


def periodic_fill(t: torch.Tensor, m: int) -> torch.Tensor:
    """bdperx then bdpery on a (jx, ix) tensor (returns a new tensor)."""
    t = t.clone()
    ix, jx = t.shape[1], t.shape[0]
    t[:, :m] = t[:, ix - 2 * m: ix - m]
    t[:, ix - m:] = t[:, m: 2 * m]
    t[:m, :] = t[jx - 2 * m: jx - m, :]
    t[jx - m:, :] = t[m: 2 * m, :]
    return t


class FinalDensity(torch.autograd.Function):
    """rho after ``solver.nsteps`` steps (interior) as a function of the
    initial interior velocities vx0, vy0 and a uniform pressure scale beta
    (p0 = beta * base pressure); the other initial fields come from
    ``base`` (7 tensors with ghosts filled)."""

    @staticmethod
    def forward(ctx, vx0, vy0, beta, solver, base):
        m = solver.margin
        vx = base[2].clone()
        vy = base[3].clone()
        vx[m:-m, m:-m] = vx0
        vy[m:-m, m:-m] = vy0
        state = [base[0], base[1] * beta, periodic_fill(vx, m), periodic_fill(vy, m), base[4], base[5], base[6]]
        final = solver.forward(state)
        ctx.solver, ctx.base = solver, base
        ctx.beta_shape = beta.shape
        return final[0][m:-m, m:-m].clone()

    @staticmethod
    def backward(ctx, g):
        solver, base = ctx.solver, ctx.base
        m = solver.margin
        lam = [torch.zeros_like(base[0]) for _ in range(6)]
        lam[0][m:-m, m:-m] = g
        l0 = solver.backward(lam)
        gvx = l0[2][m:-m, m:-m].clone()
        gvy = l0[3][m:-m, m:-m].clone()
        gbeta = (l0[1][m:-m, m:-m] * base[1][m:-m, m:-m]).sum().reshape(ctx.beta_shape)
        return gvx, gvy, gbeta, None, None
