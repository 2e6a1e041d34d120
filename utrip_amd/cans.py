"""Python mirror of the CANS hot-path routines over libcans_gpu.so.

Every function takes torch float64 device tensors laid out as the reference
holds its arrays: a Fortran array ``a(ix,jx,kx)`` is a contiguous tensor of
shape ``(kx, jx, ix)`` (1D: ``(ix,)``, 2D: ``(jx, ix)``), so element
``(i,j,k)`` (1-based) is ``t[k-1, j-1, i-1]``.  Names and argument order follow
the reference routines (``cans1d/hdmlw/mlw_m.f:2`` etc.); arrays are updated
in place, scalars are Python floats.  Errors raise :class:`CansError`; the
library must be built (``python -m utrip_amd.build``) -- there is no fallback.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import torch

from . import LIB_GPU

BD_FRE, BD_SPP, BD_SPN, BD_SMP, BD_SMN, BD_CNS, BD_FRD, BD_PER, BD_NONE = 0, 1, 2, 3, 4, 5, 6, 7, -1
BND_NAMES = {"fre": BD_FRE, "spp": BD_SPP, "spn": BD_SPN, "smp": BD_SMP, "smn": BD_SMN,
             "cns": BD_CNS, "frd": BD_FRD, "per": BD_PER}


class CansError(RuntimeError):
    pass


_lib: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    """Load libcans_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_GPU.exists():
            raise CansError(f"{LIB_GPU} is not built; run `python -m utrip_amd.build`")
        _lib = C.CDLL(str(LIB_GPU))
        _lib.cans_gpu_last_error.restype = C.c_char_p
        _lib.cans_gpu_map.restype = C.c_void_p
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise CansError(f"libcans_gpu error {rc}: {lib().cans_gpu_last_error().decode()}")


def set_march_waves(waves: int) -> None:
    """Minimum wavefronts per launch of the marching face kernels (HLLD y/z
    sweeps, 3D Roe y/z fluxes; default 32768): fewer waves, longer segments.
    Results do not depend on it."""
    _check(lib().cans_gpu_set_march_waves(int(waves)))


def set_mlw3d_path(path: int) -> None:
    """Cartesian 3D MLW step: 1 (default) one kernel marching in z with its
    cell and corner planes in LDS, 0 three kernels over HBM scratch.  Both
    give the same bits."""
    _check(lib().cans_gpu_set_mlw3d_path(int(path)))


def _stream() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _d(x: float) -> C.c_double:
    return C.c_double(float(x))


def _ptr(t: Optional[torch.Tensor], shape=None, name="array") -> C.c_void_p:
    if t is None:
        return C.c_void_p(0)
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch.Tensor")
    if t.dtype != torch.float64:
        raise TypeError(f"{name}: dtype {t.dtype}, expected float64")
    if not t.is_cuda:
        raise ValueError(f"{name}: tensor is not on a GPU")
    if not t.is_contiguous():
        raise ValueError(f"{name}: tensor is not contiguous")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(t.shape)}, expected {tuple(shape)}")
    return C.c_void_p(t.data_ptr())


def _dims(t: torch.Tensor):
    """(ix, jx, kx) of a reference-layout tensor."""
    s = tuple(t.shape)
    if len(s) == 1:
        return s[0], 1, 1
    if len(s) == 2:
        return s[1], s[0], 1
    return s[2], s[1], s[0]


# ----------------------------------------------------------------------- 1D
def mlw_m_1d(ro, pr, vx, vy, by, bx, bxm, dt, qav, gm, dx, dxm):
    """cans1d/hdmlw/mlw_m.f:2  (in place on i=2..ix-1)."""
    ix = ro.shape[0]
    sh = (ix,)
    _check(lib().cans1d_mlw_m(*[_ptr(t, sh, n) for t, n in
                                zip((ro, pr, vx, vy, by, bx, bxm), ("ro", "pr", "vx", "vy", "by", "bx", "bxm"))],
                              _d(dt), _d(qav), _d(gm), _ptr(dx, sh, "dx"), _ptr(dxm, sh, "dxm"), C.c_int(ix),
                              _stream()))


NOALLOC_MAX_IX = 65536   # CANS_NOALLOC_MAX_IX of include/cans_gpu.h


def mlw_m_1d_noalloc(ro, pr, vx, vy, by, bx, bxm, dt, qav, gm, dx, dxm):
    """cans1d/hdmlw/mlw_m.f:2 as mlw_m_1d, through the entry point that
    allocates no device memory (workspace: a static device array, so
    ix <= NOALLOC_MAX_IX; calls that overlap on one GPU race)."""
    ix = ro.shape[0]
    sh = (ix,)
    _check(lib().cans1d_mlw_m_noalloc(*[_ptr(t, sh, n) for t, n in
                                        zip((ro, pr, vx, vy, by, bx, bxm),
                                            ("ro", "pr", "vx", "vy", "by", "bx", "bxm"))],
                                      _d(dt), _d(qav), _d(gm), _ptr(dx, sh, "dx"), _ptr(dxm, sh, "dxm"),
                                      C.c_int(ix), _stream()))


def roe_m_1d(ro, pr, vx, vy, by, bx, bxm, dt, gm, dx):
    """cans1d/hdroe/roe_m.f:2  (in place on i=3..ix-2)."""
    ix = ro.shape[0]
    sh = (ix,)
    _check(lib().cans1d_roe_m(*[_ptr(t, sh) for t in (ro, pr, vx, vy, by, bx, bxm)],
                              _d(dt), _d(gm), _ptr(dx, sh), C.c_int(ix), _stream()))


def cfl_m_1d(safety, dtmin, gm, bx, ro, pr, vx, vy, by, dx):
    """cans1d/common/cfl_m.f:2  -> (dt, merr)."""
    ix = ro.shape[0]
    sh = (ix,)
    dt = C.c_double()
    merr = C.c_int()
    _check(lib().cans1d_cfl_m(C.byref(dt), _d(safety), _d(dtmin), C.byref(merr), _d(gm),
                              *[_ptr(t, sh) for t in (bx, ro, pr, vx, vy, by, dx)], C.c_int(ix), _stream()))
    return dt.value, merr.value


def bnd_1d(kind, mbnd, margin, qq, q0=0.0, dxm=None):
    """cans1d/bc/bd*x.f  (kind: CANS_BD_* or a name such as 'fre', 'per')."""
    kind = BND_NAMES.get(kind, kind)
    ix = qq.shape[0]
    _check(lib().cans1d_bnd(C.c_int(kind), C.c_int(mbnd), C.c_int(margin), _ptr(qq, (ix,)), _d(q0),
                            _ptr(dxm, (ix,)) if dxm is not None else C.c_void_p(0), C.c_int(ix), _stream()))


# ----------------------------------------------------------------------- 2D
def mlw_m_2d(ro, pr, vx, vy, bx, by, az, dt, qav, gm, dx, dxm, dy, dym):
    """cans2d/hdmlw/mlw_m.f:2-3  (in place on i=2..ix-1, j=2..jx-1)."""
    jx, ix = ro.shape
    sh = (jx, ix)
    _check(lib().cans2d_mlw_m(*[_ptr(t, sh) for t in (ro, pr, vx, vy, bx, by, az)],
                              _d(dt), _d(qav), _d(gm), _ptr(dx, (ix,)), _ptr(dxm, (ix,)), C.c_int(ix),
                              _ptr(dy, (jx,)), _ptr(dym, (jx,)), C.c_int(jx), _stream()))


def roe_m_2d(ro, pr, vx, vy, bx, by, az, dt, gm, dx, dy):
    """cans2d/hdroe/roe_m.f:2  (in place on i=3..ix-2, j=3..jx-2)."""
    jx, ix = ro.shape
    sh = (jx, ix)
    _check(lib().cans2d_roe_m(*[_ptr(t, sh) for t in (ro, pr, vx, vy, bx, by, az)],
                              _d(dt), _d(gm), _ptr(dx, (ix,)), C.c_int(ix), _ptr(dy, (jx,)), C.c_int(jx),
                              _stream()))


def cfl_m_2d(safety, dtmin, gm, ro, pr, vx, vy, bx, by, dx, dy):
    """cans2d/common/cfl_m.f:2-3  -> (dt, merr)."""
    jx, ix = ro.shape
    sh = (jx, ix)
    dt = C.c_double()
    merr = C.c_int()
    _check(lib().cans2d_cfl_m(C.byref(dt), _d(safety), _d(dtmin), C.byref(merr), _d(gm),
                              *[_ptr(t, sh) for t in (ro, pr, vx, vy, bx, by)],
                              _ptr(dx, (ix,)), C.c_int(ix), _ptr(dy, (jx,)), C.c_int(jx), _stream()))
    return dt.value, merr.value


def bnd_2d(dir, kind, mbnd, margin, qq, q0=0.0, dm=None):
    """cans2d/bc/bd*{x,y}.f  (dir 0 = x, 1 = y)."""
    kind = BND_NAMES.get(kind, kind)
    jx, ix = qq.shape
    n = ix if dir == 0 else jx
    _check(lib().cans2d_bnd(C.c_int(dir), C.c_int(kind), C.c_int(mbnd), C.c_int(margin), _ptr(qq, (jx, ix)),
                            _d(q0), _ptr(dm, (n,)) if dm is not None else C.c_void_p(0),
                            C.c_int(ix), C.c_int(jx), _stream()))


# ----------------------------------------------------------------------- 3D
def mlw_m_3d(ro, pr, vx, vy, vz, bx, by, bz, dt, qav, gm, dx, dxm, dy, dym, dz, dzm):
    """cans3d/hdmlw/mlw_m.f:2-3  (in place on the interior)."""
    kx, jx, ix = ro.shape
    sh = (kx, jx, ix)
    _check(lib().cans3d_mlw_m(*[_ptr(t, sh) for t in (ro, pr, vx, vy, vz, bx, by, bz)],
                              _d(dt), _d(qav), _d(gm), _ptr(dx, (ix,)), _ptr(dxm, (ix,)), C.c_int(ix),
                              _ptr(dy, (jx,)), _ptr(dym, (jx,)), C.c_int(jx),
                              _ptr(dz, (kx,)), _ptr(dzm, (kx,)), C.c_int(kx), _stream()))


def mlw_m_c_3d(ro, pr, vx, vy, vz, bx, by, bz, dt, qav, gm, x, xm, dx, dxm, dy, dym, dz, dzm):
    """cans3d/hdmlw/mlw_m_c.f:2-3, cylindrical (r, phi, z)  (in place on the interior)."""
    kx, jx, ix = ro.shape
    sh = (kx, jx, ix)
    _check(lib().cans3d_mlw_m_c(*[_ptr(t, sh) for t in (ro, pr, vx, vy, vz, bx, by, bz)],
                                _d(dt), _d(qav), _d(gm), _ptr(x, (ix,)), _ptr(xm, (ix,)),
                                _ptr(dx, (ix,)), _ptr(dxm, (ix,)), C.c_int(ix),
                                _ptr(dy, (jx,)), _ptr(dym, (jx,)), C.c_int(jx),
                                _ptr(dz, (kx,)), _ptr(dzm, (kx,)), C.c_int(kx), _stream()))


def mlw_m_s_3d(ro, pr, vx, vy, vz, bx, by, bz, dt, qav, gm, x, xm, dx, dxm, y, ym, dy, dym, dz, dzm):
    """cans3d/hdmlw/mlw_m_s.f:2-3, spherical (r, theta, phi)  (in place on the interior)."""
    kx, jx, ix = ro.shape
    sh = (kx, jx, ix)
    _check(lib().cans3d_mlw_m_s(*[_ptr(t, sh) for t in (ro, pr, vx, vy, vz, bx, by, bz)],
                                _d(dt), _d(qav), _d(gm), _ptr(x, (ix,)), _ptr(xm, (ix,)),
                                _ptr(dx, (ix,)), _ptr(dxm, (ix,)), C.c_int(ix), _ptr(y, (jx,)), _ptr(ym, (jx,)),
                                _ptr(dy, (jx,)), _ptr(dym, (jx,)), C.c_int(jx),
                                _ptr(dz, (kx,)), _ptr(dzm, (kx,)), C.c_int(kx), _stream()))


def cfl_m_s_3d(safety, dtmin, gm, ro, pr, vx, vy, vz, bx, by, bz, x, dx, y, dy, dz):
    """cans3d/common/cfl_m_s.f:1-2  -> (dt, merr)."""
    kx, jx, ix = ro.shape
    sh = (kx, jx, ix)
    dt = C.c_double()
    merr = C.c_int()
    _check(lib().cans3d_cfl_m_s(C.byref(dt), _d(safety), _d(dtmin), C.byref(merr), _d(gm),
                                *[_ptr(t, sh) for t in (ro, pr, vx, vy, vz, bx, by, bz)],
                                _ptr(x, (ix,)), _ptr(dx, (ix,)), C.c_int(ix), _ptr(y, (jx,)), _ptr(dy, (jx,)),
                                C.c_int(jx), _ptr(dz, (kx,)), C.c_int(kx), _stream()))
    return dt.value, merr.value


def roe_m_3d(ro, pr, vx, vy, vz, bx, by, bz, dt, gm, dx, dy, dz):
    """cans3d/hdroe/roe_m.f:2."""
    kx, jx, ix = ro.shape
    sh = (kx, jx, ix)
    _check(lib().cans3d_roe_m(*[_ptr(t, sh) for t in (ro, pr, vx, vy, vz, bx, by, bz)],
                              _d(dt), _d(gm), _ptr(dx, (ix,)), C.c_int(ix), _ptr(dy, (jx,)), C.c_int(jx),
                              _ptr(dz, (kx,)), C.c_int(kx), _stream()))


def cfl_m_3d(safety, dtmin, gm, ro, pr, vx, vy, vz, bx, by, bz, dx, dy, dz):
    """cans3d/common/cfl_m.f:2-3  -> (dt, merr)."""
    kx, jx, ix = ro.shape
    sh = (kx, jx, ix)
    dt = C.c_double()
    merr = C.c_int()
    _check(lib().cans3d_cfl_m(C.byref(dt), _d(safety), _d(dtmin), C.byref(merr), _d(gm),
                              *[_ptr(t, sh) for t in (ro, pr, vx, vy, vz, bx, by, bz)],
                              _ptr(dx, (ix,)), C.c_int(ix), _ptr(dy, (jx,)), C.c_int(jx),
                              _ptr(dz, (kx,)), C.c_int(kx), _stream()))
    return dt.value, merr.value


def bnd_3d(dir, kind, mbnd, margin, qq, q0=0.0, dm=None):
    """cans3d/bc/bd*{x,y,z}.f  (dir 0 = x, 1 = y, 2 = z)."""
    kind = BND_NAMES.get(kind, kind)
    kx, jx, ix = qq.shape
    n = (ix, jx, kx)[dir]
    _check(lib().cans3d_bnd(C.c_int(dir), C.c_int(kind), C.c_int(mbnd), C.c_int(margin),
                            _ptr(qq, (kx, jx, ix)), _d(q0),
                            _ptr(dm, (n,)) if dm is not None else C.c_void_p(0),
                            C.c_int(ix), C.c_int(jx), C.c_int(kx), _stream()))


def hlld_engine_3d(ro, pr, vx, vy, vz, bx, by, bz, gm, mu, dt, dx, dy, dz, mfdim=(1, 1, 1)):
    """cansxx-nosupport/md_mhdshktb/hdhlld-x/engine.f:2  (HLLD MUSCL-Hancock)."""
    kx, jx, ix = ro.shape
    sh = (kx, jx, ix)
    mf = (C.c_int * 3)(*[int(m) for m in mfdim])
    _check(lib().cans3d_hlld_engine(*[_ptr(t, sh) for t in (ro, pr, vx, vy, vz, bx, by, bz)],
                                    _d(gm), _d(mu), _d(dt), _ptr(dx, (ix,)), _ptr(dy, (jx,)), _ptr(dz, (kx,)),
                                    C.c_int(ix), C.c_int(jx), C.c_int(kx), mf, _stream()))


# ----------------------------------------------------------------------- resident loop
class Clock:
    """Device clock of a resident time loop (t, ns, merr, dt on the GPU)."""

    def __init__(self):
        h = C.c_void_p()
        _check(lib().cans_clock_create(C.byref(h)))
        self.handle = h
        self.issued = 0

    def reset(self, t0=0.0, tend=1e300, nstop=2**31 - 1, dtmin=1e-10):
        _check(lib().cans_clock_reset(self.handle, _d(t0), _d(tend), C.c_int(nstop), _d(dtmin), _stream()))
        self.issued = 0

    def read(self):
        t, dtl = C.c_double(), C.c_double()
        ns, nexec, merr = C.c_int(), C.c_int(), C.c_int()
        _check(lib().cans_clock_read(self.handle, C.c_int(self.issued), C.byref(t), C.byref(ns), C.byref(nexec),
                                     C.byref(merr), C.byref(dtl), _stream()))
        return {"t": t.value, "ns": ns.value, "nexec": nexec.value, "merr": merr.value, "dt": dtl.value}

    def __del__(self):
        try:
            if self.handle and _lib is not None:
                _lib.cans_clock_destroy(self.handle)
        except Exception:
            pass


def _ptr_array(ts: Sequence[torch.Tensor], shape):
    arr = (C.c_void_p * len(ts))()
    for n, t in enumerate(ts):
        arr[n] = _ptr(t, shape).value
    return arr


class MlwRun2D:
    """Resident 2D MLW time loop: {cfl_m; t=t+dt; mlw_m; bnd} back to back.

    ``state`` is a list of 7 tensors (ro,pr,vx,vy,bx,by,az) with ghosts
    filled; a second buffer set is allocated.  ``bc=None`` means a periodic
    box (ghost fill and the next step's cfl fused into the engine kernel).
    """

    def __init__(self, state, dx, dxm, dy, dym, qav, gm, safety, margin=1, bc=None, clock=None):
        self.a = list(state)
        self.b = [torch.empty_like(t) for t in state]
        self.jx, self.ix = state[0].shape
        self.grid = (dx, dxm, dy, dym)
        self.qav, self.gm, self.safety, self.margin = qav, gm, safety, margin
        self.clock = clock or Clock()
        n = lib().cans2d_metric_ws_size(self.ix, self.jx)
        self.metric = torch.empty(n, dtype=torch.float64, device=state[0].device)
        if bc is None:
            self.bc = None
        else:
            self.bc = (C.c_int * 28)(*[int(k) for k in bc])
        self._pa = _ptr_array(self.a, (self.jx, self.ix))
        self._pb = _ptr_array(self.b, (self.jx, self.ix))

    def reset(self, t0=0.0, tend=1e300, nstop=2**31 - 1, dtmin=1e-10):
        self.clock.reset(t0, tend, nstop, dtmin)

    def steps(self, n: int, events=None):
        """Issue n steps; `events` (an EventSet of >= 2n events) brackets the
        engine kernel of every step."""
        dx, dxm, dy, dym = self.grid
        _check(lib().cans2d_mlw_run(self._pa, self._pb, self.clock.handle, C.c_int(self.clock.issued), C.c_int(n),
                                    _d(self.qav), _d(self.gm), _d(self.safety),
                                    _ptr(dx, (self.ix,)), _ptr(dxm, (self.ix,)), C.c_int(self.ix),
                                    _ptr(dy, (self.jx,)), _ptr(dym, (self.jx,)), C.c_int(self.jx),
                                    self.bc if self.bc is not None else None, C.c_int(self.margin),
                                    _ptr(self.metric), _events(events), _stream()))
        self.clock.issued += n

    def state(self):
        """Current state (the buffer set the last executed step wrote)."""
        info = self.clock.read()
        return (self.a if info["nexec"] % 2 == 0 else self.b), info


def _events(ev):
    """EventSet, EventSet.at(k) pointer, or None -> the C array argument."""
    if ev is None:
        return None
    return ev.array if isinstance(ev, EventSet) else ev


class EventSet:
    """n HIP events (for bracketing kernels inside a run)."""

    def __init__(self, n: int):
        self.array = (C.c_void_p * n)()
        for k in range(n):
            h = C.c_void_p()
            _check(lib().cans_event_create(C.byref(h)))
            self.array[k] = h.value
        self.n = n

    def at(self, k: int):
        """The array from event k on (a run given it brackets its steps with
        events k, k+1, ...)."""
        if not 0 <= k < self.n:
            raise IndexError(k)
        return C.cast(C.byref(self.array, k * C.sizeof(C.c_void_p)), C.POINTER(C.c_void_p))

    def record(self, k: int):
        _check(lib().cans_event_record(C.c_void_p(self.array[k]), _stream()))

    def elapsed_ms(self, a: int, b: int) -> float:
        ms = C.c_float()
        _check(lib().cans_event_elapsed_ms(C.c_void_p(self.array[a]), C.c_void_p(self.array[b]), C.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            if _lib is not None:
                for k in range(self.n):
                    _lib.cans_event_destroy(C.c_void_p(self.array[k]))
        except Exception:
            pass


def roe_m_3d_mixed(ro, pr, vx, vy, vz, bx, by, bz, dt, gm, dx, dy, dz):
    """roe_m (3D) with FP32 face-flux arithmetic, FP64 state (no reference counterpart)."""
    kx, jx, ix = ro.shape
    sh = (kx, jx, ix)
    _check(lib().cans3d_roe_m_mixed(*[_ptr(t, sh) for t in (ro, pr, vx, vy, vz, bx, by, bz)],
                                    _d(dt), _d(gm), _ptr(dx, (ix,)), C.c_int(ix), _ptr(dy, (jx,)), C.c_int(jx),
                                    _ptr(dz, (kx,)), C.c_int(kx), _stream()))


def hlld_engine_3d_mixed(ro, pr, vx, vy, vz, bx, by, bz, gm, mu, dt, dx, dy, dz, mfdim=(1, 1, 1)):
    """HLLD engine with FP32 face-flux arithmetic, FP64 state (no reference counterpart)."""
    kx, jx, ix = ro.shape
    sh = (kx, jx, ix)
    mf = (C.c_int * 3)(*[int(m) for m in mfdim])
    _check(lib().cans3d_hlld_engine_mixed(*[_ptr(t, sh) for t in (ro, pr, vx, vy, vz, bx, by, bz)],
                                          _d(gm), _d(mu), _d(dt), _ptr(dx, (ix,)), _ptr(dy, (jx,)), _ptr(dz, (kx,)),
                                          C.c_int(ix), C.c_int(jx), C.c_int(kx), mf, _stream()))


ENGINE_MLW, ENGINE_ROE, ENGINE_HLLD, ENGINE_MLW3, ENGINE_ROE3, ENGINE_MLWC, ENGINE_MLWS = 0, 1, 2, 3, 4, 5, 6
FIELDS = {1: ("ro", "pr", "vx", "vy", "by"), 2: ("ro", "pr", "vx", "vy", "bx", "by", "az"),
          3: ("ro", "pr", "vx", "vy", "vz", "bx", "by", "bz")}
# fields of the three-component engines mlw_m3 / roe_m3 (ENGINE_MLW3, ENGINE_ROE3)
FIELDS_M3 = {1: ("ro", "pr", "vx", "vy", "vz", "by", "bz"), 2: ("ro", "pr", "vx", "vy", "vz", "bx", "by", "bz", "az")}


# ----------------------------------------------------------------------- three-component variants
def mlw_m3_1d(ro, pr, vx, vy, vz, by, bz, bx, bxm, dt, qav, gm, dx, dxm):
    """cans1d/hdmlw/mlw_m3.f:2  (in place on i=2..ix-1)."""
    ix = ro.shape[0]
    sh = (ix,)
    _check(lib().cans1d_mlw_m3(*[_ptr(t, sh) for t in (ro, pr, vx, vy, vz, by, bz, bx, bxm)], _d(dt), _d(qav),
                               _d(gm), _ptr(dx, sh), _ptr(dxm, sh), C.c_int(ix), _stream()))


def cfl_m3_1d(safety, dtmin, gm, bx, ro, pr, vx, vy, by, vz, bz, dx):
    """cans1d/common/cfl_m3.f:2  -> (dt, merr)."""
    ix = ro.shape[0]
    sh = (ix,)
    dt, merr = C.c_double(), C.c_int()
    _check(lib().cans1d_cfl_m3(C.byref(dt), _d(safety), _d(dtmin), C.byref(merr), _d(gm),
                               *[_ptr(t, sh) for t in (bx, ro, pr, vx, vy, by, vz, bz, dx)], C.c_int(ix), _stream()))
    return dt.value, merr.value


def roe_m3_1d(ro, pr, vx, vy, vz, by, bz, bx, bxm, dt, gm, dx):
    """cans1d/hdroe/roe_m3.f:2  (in place on i=3..ix-2)."""
    ix = ro.shape[0]
    sh = (ix,)
    _check(lib().cans1d_roe_m3(*[_ptr(t, sh) for t in (ro, pr, vx, vy, vz, by, bz, bx, bxm)], _d(dt), _d(gm),
                               _ptr(dx, sh), C.c_int(ix), _stream()))


def roe_m3_2d(ro, pr, vx, vy, vz, bx, by, bz, az, dt, gm, dx, dy):
    """cans2d/hdroe/roe_m3.f:2  (in place on i=3..ix-2, j=3..jx-2)."""
    jx, ix = ro.shape
    sh = (jx, ix)
    _check(lib().cans2d_roe_m3(*[_ptr(t, sh) for t in (ro, pr, vx, vy, vz, bx, by, bz, az)], _d(dt), _d(gm),
                               _ptr(dx, (ix,)), C.c_int(ix), _ptr(dy, (jx,)), C.c_int(jx), _stream()))


def mlw_m_c_2d(ro, pr, vx, vz, bx, bz, ay, dt, qav, gm, x, xm, dx, dxm, dz, dzm):
    """cans2d/hdmlw/mlw_m_c.f:2-3, cylindrical (r, z)  (in place on i=2..ix-1, j=2..jx-1)."""
    jx, ix = ro.shape
    sh = (jx, ix)
    _check(lib().cans2d_mlw_m_c(*[_ptr(t, sh) for t in (ro, pr, vx, vz, bx, bz, ay)], _d(dt), _d(qav), _d(gm),
                                _ptr(x, (ix,)), _ptr(xm, (ix,)), _ptr(dx, (ix,)), _ptr(dxm, (ix,)), C.c_int(ix),
                                _ptr(dz, (jx,)), _ptr(dzm, (jx,)), C.c_int(jx), _stream()))


def mlw_m_s_2d(ro, pr, vx, vy, bx, by, az, dt, qav, gm, x, xm, dx, dxm, y, ym, dy, dym):
    """cans2d/hdmlw/mlw_m_s.f:1-2, spherical (r, theta)  (in place on i=2..ix-1, j=2..jx-1)."""
    jx, ix = ro.shape
    sh = (jx, ix)
    _check(lib().cans2d_mlw_m_s(*[_ptr(t, sh) for t in (ro, pr, vx, vy, bx, by, az)], _d(dt), _d(qav), _d(gm),
                                _ptr(x, (ix,)), _ptr(xm, (ix,)), _ptr(dx, (ix,)), _ptr(dxm, (ix,)), C.c_int(ix),
                                _ptr(y, (jx,)), _ptr(ym, (jx,)), _ptr(dy, (jx,)), _ptr(dym, (jx,)), C.c_int(jx),
                                _stream()))


def cfl_m_s_2d(safety, dtmin, gm, ro, pr, vx, vy, bx, by, x, dx, dy):
    """cans2d/common/cfl_m_s.f:1-2  -> (dt, merr)."""
    jx, ix = ro.shape
    sh = (jx, ix)
    dt, merr = C.c_double(), C.c_int()
    _check(lib().cans2d_cfl_m_s(C.byref(dt), _d(safety), _d(dtmin), C.byref(merr), _d(gm),
                                *[_ptr(t, sh) for t in (ro, pr, vx, vy, bx, by)], _ptr(x, (ix,)), _ptr(dx, (ix,)),
                                C.c_int(ix), _ptr(dy, (jx,)), C.c_int(jx), _stream()))
    return dt.value, merr.value


def mlw_m3_2d(ro, pr, vx, vy, vz, bx, by, bz, az, dt, qav, gm, dx, dxm, dy, dym):
    """cans2d/hdmlw/mlw_m3.f:2-3  (in place on i=2..ix-1, j=2..jx-1)."""
    jx, ix = ro.shape
    sh = (jx, ix)
    _check(lib().cans2d_mlw_m3(*[_ptr(t, sh) for t in (ro, pr, vx, vy, vz, bx, by, bz, az)], _d(dt), _d(qav),
                               _d(gm), _ptr(dx, (ix,)), _ptr(dxm, (ix,)), C.c_int(ix), _ptr(dy, (jx,)),
                               _ptr(dym, (jx,)), C.c_int(jx), _stream()))


def cfl_m3_2d(safety, dtmin, gm, ro, pr, vx, vy, vz, bx, by, bz, dx, dy):
    """cans2d/common/cfl_m3.f:2-3  -> (dt, merr)."""
    jx, ix = ro.shape
    sh = (jx, ix)
    dt, merr = C.c_double(), C.c_int()
    _check(lib().cans2d_cfl_m3(C.byref(dt), _d(safety), _d(dtmin), C.byref(merr), _d(gm),
                               *[_ptr(t, sh) for t in (ro, pr, vx, vy, vz, bx, by, bz)],
                               _ptr(dx, (ix,)), C.c_int(ix), _ptr(dy, (jx,)), C.c_int(jx), _stream()))
    return dt.value, merr.value


class Run:
    """Resident CANS time loop for any dimension and engine (cans_run).

    ``state``: list of the dimension's fields (see FIELDS) with ghosts filled.
    ``grid``: dict with dx, dxm (and dy, dym, dz, dzm as the dimension needs).
    ``bc``: per (dir, side, field) CANS_BD_* kinds, a flat list of
    2*dim*nf ints, or a single kind for every face and field; None = the
    fused periodic 2D MLW path.
    """

    def __init__(self, dim, engine, state, grid, gm, safety, qav=3.0, mu=1.0, margin=1, bc=None, mixed=False,
                 bx1d=None, bxm1d=None, clock=None):
        self.dim, self.engine, self.mixed = dim, engine, mixed
        self.nf = len(FIELDS_M3[dim] if engine in (ENGINE_MLW3, ENGINE_ROE3) else FIELDS[dim])
        assert len(state) == self.nf, f"{dim}D needs {self.nf} fields"
        self.a = list(state)
        self.b = [torch.empty_like(t) for t in state]
        shape = tuple(state[0].shape)
        self.ix, self.jx, self.kx = _dims(state[0])
        self.shape = shape
        self.grid = grid
        self.gm, self.safety, self.qav, self.mu, self.margin = gm, safety, qav, mu, margin
        self.bx1d, self.bxm1d = bx1d, bxm1d
        self.clock = clock or Clock()
        lib().cans_run_ws_bytes.restype = C.c_size_t
        nbytes = lib().cans_run_ws_bytes(C.c_int(dim), C.c_int(engine), C.c_int(self.ix), C.c_int(self.jx),
                                         C.c_int(self.kx))
        self.ws = torch.empty((nbytes + 7) // 8, dtype=torch.float64, device=state[0].device)
        if bc is None:
            self.bc = None
        else:
            if isinstance(bc, int):
                bc = [bc] * (2 * dim * self.nf)
            assert len(bc) == 2 * dim * self.nf
            self.bc = (C.c_int * len(bc))(*[int(k) for k in bc])
        self._pa = _ptr_array(self.a, shape)
        self._pb = _ptr_array(self.b, shape)

    def reset(self, t0=0.0, tend=1e300, nstop=2**31 - 1, dtmin=1e-10):
        self.clock.reset(t0, tend, nstop, dtmin)

    def steps(self, n: int, events=None):
        g = self.grid
        # bx1d: 1D bx, or the radii x (cylindrical), or x, xm, y, ym back to back (spherical)
        nb1 = 2 * (self.ix + self.jx) if self.engine == ENGINE_MLWS else self.ix

        def gp(name, ln):
            return _ptr(g[name], (ln,)) if name in g and g[name] is not None else C.c_void_p(0)
        _check(lib().cans_run(C.c_int(self.dim), C.c_int(self.engine), C.c_int(1 if self.mixed else 0),
                              self._pa, self._pb,
                              _ptr(self.bx1d, (nb1,)) if self.bx1d is not None else C.c_void_p(0),
                              _ptr(self.bxm1d, (self.ix,)) if self.bxm1d is not None else C.c_void_p(0),
                              self.clock.handle, C.c_int(self.clock.issued), C.c_int(n),
                              _d(self.qav), _d(self.gm), _d(self.mu), _d(self.safety),
                              gp("dx", self.ix), gp("dxm", self.ix), C.c_int(self.ix),
                              gp("dy", self.jx), gp("dym", self.jx), C.c_int(self.jx),
                              gp("dz", self.kx), gp("dzm", self.kx), C.c_int(self.kx),
                              self.bc, C.c_int(self.margin), _ptr(self.ws),
                              _events(events), _stream()))
        self.clock.issued += n

    def state(self):
        info = self.clock.read()
        return (self.a if info["nexec"] % 2 == 0 else self.b), info


def _iptr(t: torch.Tensor, dtype, n: int, name: str) -> C.c_void_p:
    if t.dtype != dtype or not t.is_cuda or not t.is_contiguous() or t.numel() != n:
        raise ValueError(f"{name}: expected a contiguous {dtype} GPU tensor of {n} entries")
    return C.c_void_p(t.data_ptr())


class Batch1D:
    """Many independent 1D tubes, each running the cans1d main loop
    {cfl_m; t=t+dt; mlw_m; bdfrex both ends} to its own end
    (cans1d/md_mhdshktb/hdmlw/main.f:132-186), in one persistent kernel
    (cans1d_batch_run).

    ``tubes``: a list of dicts of 1D numpy arrays ro, pr, vx, vy, by, bx,
    bxm, dx, dxm (ghosts included, margin filled).  The arrays are packed
    into one device array per field; tube k occupies ``slice(off[k],
    off[k]+ix[k])``.  ``mixed``: FP32 fluxes / half step, FP64 accumulation.
    The queue order is longest-first by the estimated step count
    ``ix / dt0`` (dt0 from the initial cfl_m)."""

    FIELDS = ("ro", "pr", "vx", "vy", "by")
    COEF = ("bx", "bxm", "dx", "dxm")

    def __init__(self, tubes, gm, qav, safety, margin, dtmin=1e-10, mixed=False, device="cuda:0", tune_room=True):
        import numpy as np
        self.n = len(tubes)
        self.gm, self.qav, self.safety, self.margin, self.dtmin, self.mixed = gm, qav, safety, margin, dtmin, mixed
        ix = np.array([len(t["ro"]) for t in tubes], dtype=np.int32)
        off = np.zeros(self.n, dtype=np.int64)
        if self.n:
            off[1:] = np.cumsum(ix[:-1], dtype=np.int64)
        self.ncells = int(ix.sum())
        self.ix_host, self.off_host = ix, off
        dev = torch.device(device)

        def pack(k):
            a = np.concatenate([np.asarray(t[k], dtype=np.float64) for t in tubes]) if self.n else np.zeros(0)
            return torch.from_numpy(a).to(dev)

        self.f = [pack(k) for k in self.FIELDS]
        self.c = [pack(k) for k in self.COEF]
        # longest first: steps ~ 1/dt0, work per step ~ ix
        cost = np.zeros(self.n)
        for k, t in enumerate(tubes):
            ro, pr, vx, by, bx, dx = (np.asarray(t[q], dtype=np.float64) for q in ("ro", "pr", "vx", "by", "bx", "dx"))
            s = slice(1, len(ro) - 1)
            cf = np.abs(vx[s]) + np.sqrt(gm * pr[s] / ro[s] + (bx[s] ** 2 + by[s] ** 2) * 0.25 / np.pi / ro[s])
            cost[k] = len(ro) * float(np.max(cf / dx[s]))
        self.order_host = np.argsort(-cost, kind="stable").astype(np.int32)
        self.off = torch.from_numpy(off).to(dev)
        self.ix = torch.from_numpy(ix).to(dev)
        self.order = torch.from_numpy(self.order_host).to(dev)
        self.t = torch.zeros(self.n, dtype=torch.float64, device=dev)
        self.ns = torch.zeros(self.n, dtype=torch.int32, device=dev)
        self.merr = torch.zeros(self.n, dtype=torch.int32, device=dev)
        lib().cans1d_batch_ws_bytes.restype = C.c_size_t
        nb = lib().cans1d_batch_ws_bytes(C.c_longlong(self.ncells), C.c_int(1 if tune_room else 0))
        self.ws = torch.empty((nb + 7) // 8, dtype=torch.float64, device=dev)
        self.shape = (256, 2)   # (block_threads, blocks_per_cu) until tuned
        self._graphs = []
        torch.cuda.synchronize()   # off/ix/order are read on the host by the calls

    def _common(self):
        n, nc = self.n, self.ncells
        return [C.c_int(n), _iptr(self.off, torch.int64, n, "off"), _iptr(self.ix, torch.int32, n, "ix"),
                _iptr(self.order, torch.int32, n, "order"), *[_ptr(t, (nc,)) for t in self.f],
                *[_ptr(t, (nc,)) for t in self.c], C.c_longlong(nc), _d(self.gm), _d(self.qav), _d(self.safety),
                _d(self.dtmin)]

    def _outs(self):
        n = self.n
        return [_ptr(self.t, (n,)), _iptr(self.ns, torch.int32, n, "ns"), _iptr(self.merr, torch.int32, n, "merr")]

    def tune(self, probe_steps: int = 20):
        """Pick (block_threads, blocks_per_cu) by timing every candidate on
        a copy of the state (the state is unchanged)."""
        bt, bpc, ms = C.c_int(), C.c_int(), C.c_float()
        _check(lib().cans1d_batch_tune(*self._common(), C.c_int(self.margin), C.c_int(1 if self.mixed else 0),
                                       C.c_int(probe_steps), *self._outs(), _ptr(self.ws), C.byref(bt), C.byref(bpc),
                                       C.byref(ms), _stream()))
        self.shape = (bt.value, bpc.value)
        return bt.value, bpc.value, ms.value

    def run(self, tend: float, nstop: int = 2**31 - 1):
        bt, bpc = self.shape
        _check(lib().cans1d_batch_run(*self._common(), _d(tend), C.c_int(nstop), C.c_int(self.margin),
                                      C.c_int(1 if self.mixed else 0), *self._outs(), C.c_int(bt), C.c_int(bpc),
                                      _ptr(self.ws), _stream()))

    def run_steps(self, nstop_tube: torch.Tensor, tend: float = 1e300):
        """Run tube k while ns < nstop_tube[k] and t < tend (nstop_tube: an
        int32 device tensor of one entry per tube)."""
        bt, bpc = self.shape
        _check(lib().cans1d_batch_run_steps(*self._common(), _d(tend),
                                            _iptr(nstop_tube, torch.int32, self.n, "nstop_tube"),
                                            C.c_int(self.margin), C.c_int(1 if self.mixed else 0), *self._outs(),
                                            C.c_int(bt), C.c_int(bpc), _ptr(self.ws), _stream()))

    def graph(self, tend: float, nstop: int = 2**31 - 1, init=None):
        """An instantiated hipGraph of the whole run; ``init`` (5 packed
        device tensors ro..by) is copied into the state at the start of
        every replay."""
        bt, bpc = self.shape
        arr = None
        if init is not None:
            arr = (C.c_void_p * 5)(*[_ptr(t, (self.ncells,)).value for t in init])
        h = C.c_void_p()
        _check(lib().cans1d_batch_graph(*self._common(), _d(tend), C.c_int(nstop), C.c_int(self.margin),
                                        C.c_int(1 if self.mixed else 0), *self._outs(), C.c_int(bt), C.c_int(bpc),
                                        _ptr(self.ws), arr, C.byref(h)))
        g = BatchGraph(h)
        self._graphs.append(g)
        return g

    def tube(self, k: int):
        """Host copy of tube k's fields."""
        s = slice(int(self.off_host[k]), int(self.off_host[k] + self.ix_host[k]))
        return {q: t[s].cpu().numpy() for q, t in zip(self.FIELDS, self.f)}


class Batch2D:
    """Many independent periodic 2D MLW runs, each running the cans2d main
    loop {cfl_m; t=t+dt; mlw_m; bdperx; bdpery} to its own end
    (cans2d/md_mhdshktb/hdmlw/main.f:150-206), in one persistent kernel
    (cans2d_batch_run).

    ``runs``: a list of dicts of Fortran-ordered numpy arrays ro, pr, vx, vy,
    bx, by, az of shape (ix, jx) (ghosts included and filled periodically)
    and dx, dxm (ix), dy, dym (jx).  Run k's fields occupy cells
    ``off[k] .. off[k] + ix*jx - 1`` of one packed device array per field.
    The queue order is costliest first (cells / dt0, dt0 from the initial
    cfl_m)."""

    FIELDS = ("ro", "pr", "vx", "vy", "bx", "by", "az")

    def __init__(self, runs, gm, qav, safety, margin, dtmin=1e-10, device="cuda:0", tune_room=True):
        import numpy as np
        self.n = len(runs)
        self.gm, self.qav, self.safety, self.margin, self.dtmin = gm, qav, safety, margin, dtmin
        ix = np.array([r["ro"].shape[0] for r in runs], dtype=np.int32)
        jx = np.array([r["ro"].shape[1] for r in runs], dtype=np.int32)
        cells = ix.astype(np.int64) * jx

        def starts(sizes):
            o = np.zeros(self.n, dtype=np.int64)
            if self.n:
                o[1:] = np.cumsum(sizes[:-1], dtype=np.int64)
            return o

        off, offx, offy = starts(cells), starts(ix.astype(np.int64)), starts(jx.astype(np.int64))
        self.ncells, self.nx1d, self.ny1d = int(cells.sum()), int(ix.sum()), int(jx.sum())
        self.ix_host, self.jx_host, self.off_host = ix, jx, off
        dev = torch.device(device)

        def pack(get):
            a = np.concatenate([np.asarray(get(r), dtype=np.float64).ravel(order="F") for r in runs])
            return torch.from_numpy(a).to(dev)

        self.f = [pack(lambda r, k=k: r[k]) for k in self.FIELDS]
        self.g1 = [pack(lambda r, k=k: r[k]) for k in ("dx", "dxm", "dy", "dym")]
        cost = np.zeros(self.n)
        for k, r in enumerate(runs):
            ro, pr, vx, vy, bx, by = (np.asarray(r[q], dtype=np.float64) for q in self.FIELDS[:6])
            c = np.sqrt(vx ** 2 + vy ** 2 + gm * pr / ro + (bx ** 2 + by ** 2) / np.pi / 4.0 / ro)
            d = np.minimum(np.asarray(r["dx"])[:, None], np.asarray(r["dy"])[None, :])
            cost[k] = float(cells[k]) * float(np.max(c / d))
        self.order_host = np.argsort(-cost, kind="stable").astype(np.int32)
        to = lambda a: torch.from_numpy(a).to(dev)
        self.off, self.offx, self.offy = to(off), to(offx), to(offy)
        self.ix, self.jx, self.order = to(ix), to(jx), to(self.order_host)
        self.t = torch.zeros(self.n, dtype=torch.float64, device=dev)
        self.ns = torch.zeros(self.n, dtype=torch.int32, device=dev)
        self.merr = torch.zeros(self.n, dtype=torch.int32, device=dev)
        lib().cans2d_batch_ws_bytes.restype = C.c_size_t
        nb = lib().cans2d_batch_ws_bytes(C.c_longlong(self.ncells), C.c_longlong(self.nx1d), C.c_longlong(self.ny1d),
                                         C.c_int(self.n), C.c_int(1 if tune_room else 0))
        self.ws = torch.empty((nb + 7) // 8, dtype=torch.float64, device=dev)
        self.shape = (512, 1)   # (block_threads, blocks_per_cu) until tuned
        self._graphs = []
        torch.cuda.synchronize()

    def _head(self):
        n = self.n
        st = (C.c_void_p * 7)(*[_ptr(t, (self.ncells,)).value for t in self.f])
        self._st = st
        return [C.c_int(n), _iptr(self.off, torch.int64, n, "off"), _iptr(self.ix, torch.int32, n, "ix"),
                _iptr(self.jx, torch.int32, n, "jx"), _iptr(self.offx, torch.int64, n, "offx"),
                _iptr(self.offy, torch.int64, n, "offy"), _iptr(self.order, torch.int32, n, "order"), st,
                _ptr(self.g1[0], (self.nx1d,)), _ptr(self.g1[1], (self.nx1d,)), _ptr(self.g1[2], (self.ny1d,)),
                _ptr(self.g1[3], (self.ny1d,)), C.c_longlong(self.ncells), C.c_longlong(self.nx1d),
                C.c_longlong(self.ny1d), _d(self.gm), _d(self.qav), _d(self.safety), _d(self.dtmin)]

    def _outs(self):
        n = self.n
        return [_ptr(self.t, (n,)), _iptr(self.ns, torch.int32, n, "ns"), _iptr(self.merr, torch.int32, n, "merr")]

    def tune(self, probe_steps: int = 10):
        """Pick (block_threads, blocks_per_cu) by timing every candidate on a
        copy of the state (probe_steps steps of every run; the state is
        unchanged)."""
        bt, bpc, ms = C.c_int(), C.c_int(), C.c_float()
        _check(lib().cans2d_batch_tune(*self._head(), C.c_int(self.margin), C.c_int(probe_steps), *self._outs(),
                                       _ptr(self.ws), C.byref(bt), C.byref(bpc), C.byref(ms), _stream()))
        self.shape = (bt.value, bpc.value)
        return bt.value, bpc.value, ms.value

    def run(self, tend: float, nstop: int = 2**31 - 1, block_threads: int = None, blocks_per_cu: int = None):
        bt = block_threads if block_threads is not None else self.shape[0]
        bpc = blocks_per_cu if blocks_per_cu is not None else self.shape[1]
        _check(lib().cans2d_batch_run(*self._head(), _d(tend), C.c_int(nstop), C.c_int(self.margin), *self._outs(),
                                      C.c_int(bt), C.c_int(bpc), _ptr(self.ws), _stream()))

    def graph(self, tend: float, nstop: int = 2**31 - 1, init=None):
        """An instantiated hipGraph of the whole batch run; ``init`` (7
        packed device tensors ro..az) is copied into the state at the start
        of every replay."""
        bt, bpc = self.shape
        arr = None
        if init is not None:
            arr = (C.c_void_p * 7)(*[_ptr(t, (self.ncells,)).value for t in init])
        h = C.c_void_p()
        _check(lib().cans2d_batch_graph(*self._head(), _d(tend), C.c_int(nstop), C.c_int(self.margin), *self._outs(),
                                        C.c_int(bt), C.c_int(bpc), _ptr(self.ws), arr, C.byref(h)))
        g = BatchGraph(h)
        self._graphs.append(g)
        return g

    def run_fields(self, k: int):
        """Host copies (Fortran order, (ix, jx)) of run k's fields."""
        import numpy as np
        ix, jx, o = int(self.ix_host[k]), int(self.jx_host[k]), int(self.off_host[k])
        return {q: np.asfortranarray(t[o:o + ix * jx].cpu().numpy().reshape((ix, jx), order="F"))
                for q, t in zip(self.FIELDS, self.f)}


class BatchGraph:
    def __init__(self, handle):
        self.h = handle

    def launch(self):
        _check(lib().cans_graph_launch(self.h, _stream()))

    def __del__(self):
        try:
            if _lib is not None and self.h:
                _lib.cans_graph_destroy(self.h)
        except Exception:
            pass
