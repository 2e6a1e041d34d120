"""ctypes bindings of the CPU checkers (test infrastructure only).

``Oracle``  -- liboracle.so, this repository's C restatement of the reference
              routines (cans_oracle*.c).
``Ref``     -- _ref/libcans{1d,2d,3d,hlld,spicule}_ref.so, the reference Fortran
              routines compiled from their own sources (see Makefile), called
              with Fortran conventions (every argument by reference).

Arrays are numpy float64 in Fortran order with the reference's shapes
(ix,), (ix,jx) or (ix,jx,kx), so ``a[i-1, j-1]`` is the reference's
``a(i,j)``.  Calls modify the arrays in place, as the reference does.

Only tests/, the smoke check and bench.py's CPU-baseline leg use this module.
"""
from __future__ import annotations

import ctypes as C
import subprocess
import threading
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_ORACLE = HERE / "liboracle.so"
REF_DIR = HERE / "_ref"
REF_LIBS = {d: REF_DIR / f"libcans{d}_ref.so" for d in ("1d", "2d", "3d", "hlld", "spicule")}
REFDIR_SRC = Path("/root/reference/cans_beta_vary")

_P = C.POINTER(C.c_double)


def build(ref: bool = True) -> None:
    """make liboracle.so (and the reference libraries when the sources exist)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    if ref and REFDIR_SRC.exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)
        # the reference's shock-tube program, linked with the reference
        # routines and with the drop-in library (needs it built first)
        if (HERE.parent / "utrip_amd" / "libcans1d_gpu.so").exists():
            subprocess.run(["make", "-s", "-C", str(HERE), "drivers"], check=True)


def _arr(a) -> C.c_void_p:
    if a is None:
        return C.c_void_p(0)
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.f_contiguous):
        raise TypeError("expected a Fortran-ordered float64 numpy array")
    return C.c_void_p(a.ctypes.data)


def big_stack_call(fn, *args, stack_bytes=1 << 30):
    """Run fn(*args) in a thread with a large stack.

    The reference engines keep their temporaries in automatic arrays, which
    its compiler places on the stack (2D 1024^2: ~230 MB per call)."""
    out = {}

    def body():
        try:
            out["r"] = fn(*args)
        except BaseException as e:  # pragma: no cover - re-raised below
            out["e"] = e

    old = threading.stack_size()
    threading.stack_size(stack_bytes)
    try:
        t = threading.Thread(target=body)
        t.start()
        t.join()
    finally:
        threading.stack_size(old)
    if "e" in out:
        raise out["e"]
    return out.get("r")


class Oracle:
    """The C restatement (cans_oracle.h)."""

    def __init__(self, path: Path = LIB_ORACLE):
        if not path.exists():
            raise FileNotFoundError(f"{path} not built (make -C oracle)")
        self.lib = C.CDLL(str(path))

    def call(self, name, *args):
        """Call an exported oracle function by name with C conventions:
        numpy float64 arrays (Fortran order) and int32 arrays by address,
        Python floats as double, ints as int, by value."""
        cargs = []
        for a in args:
            if isinstance(a, np.ndarray):
                cargs.append(C.c_void_p(a.ctypes.data) if a.dtype == np.int32 else _arr(a))
            elif isinstance(a, bool) or isinstance(a, (int, np.integer)):
                cargs.append(C.c_int(int(a)))
            else:
                cargs.append(C.c_double(float(a)))
        getattr(self.lib, name)(*cargs)

    # ---- 1D
    def mlw_m_1d(self, ro, pr, vx, vy, by, bx, bxm, dt, qav, gm, dx, dxm):
        self.lib.or1d_mlw_m(*map(_arr, (ro, pr, vx, vy, by, bx, bxm)), C.c_double(dt), C.c_double(qav),
                            C.c_double(gm), _arr(dx), _arr(dxm), C.c_int(ro.shape[0]))

    def roe_m_1d(self, ro, pr, vx, vy, by, bx, bxm, dt, gm, dx):
        self.lib.or1d_roe_m(*map(_arr, (ro, pr, vx, vy, by, bx, bxm)), C.c_double(dt), C.c_double(gm),
                            _arr(dx), C.c_int(ro.shape[0]))

    def cfl_m_1d(self, safety, dtmin, gm, bx, ro, pr, vx, vy, by, dx):
        dt, merr, imin = C.c_double(), C.c_int(), C.c_int()
        self.lib.or1d_cfl_m(C.byref(dt), C.c_double(safety), C.c_double(dtmin), C.byref(merr), C.c_double(gm),
                            *map(_arr, (bx, ro, pr, vx, vy, by, dx)), C.c_int(ro.shape[0]), C.byref(imin))
        return dt.value, merr.value, imin.value

    def bnd_1d(self, kind, mbnd, margin, qq, q0=0.0, dxm=None):
        if kind == 7:
            self.lib.or1d_bdperx(C.c_int(mbnd), C.c_int(margin), _arr(qq), C.c_int(qq.shape[0]))
        else:
            self.lib.or1d_bd(C.c_int(kind), C.c_int(mbnd), C.c_int(margin), _arr(qq), C.c_double(q0), _arr(dxm),
                             C.c_int(qq.shape[0]))

    def grdrdy_1d(self, x):
        ix = x.shape[0]
        dx, xm, dxm = (np.zeros(ix, order="F") for _ in range(3))
        self.lib.or1d_grdrdy(_arr(dx), _arr(xm), _arr(dxm), _arr(x), C.c_int(ix))
        return dx, xm, dxm

    # ---- three-component variants (mlw_m3, cfl_m3)
    def mlw_m3_1d(self, ro, pr, vx, vy, vz, by, bz, bx, bxm, dt, qav, gm, dx, dxm):
        self.lib.or1d_mlw_m3(*map(_arr, (ro, pr, vx, vy, vz, by, bz, bx, bxm)), C.c_double(dt), C.c_double(qav),
                             C.c_double(gm), _arr(dx), _arr(dxm), C.c_int(ro.shape[0]))

    def cfl_m3_1d(self, safety, dtmin, gm, bx, ro, pr, vx, vy, by, vz, bz, dx):
        dt, merr, imin = C.c_double(), C.c_int(), C.c_int()
        self.lib.or1d_cfl_m3(C.byref(dt), C.c_double(safety), C.c_double(dtmin), C.byref(merr), C.c_double(gm),
                             *map(_arr, (bx, ro, pr, vx, vy, by, vz, bz, dx)), C.c_int(ro.shape[0]), C.byref(imin))
        return dt.value, merr.value, imin.value

    def roe_m3_1d(self, ro, pr, vx, vy, vz, by, bz, bx, bxm, dt, gm, dx):
        self.lib.or1d_roe_m3(*map(_arr, (ro, pr, vx, vy, vz, by, bz, bx, bxm)), C.c_double(dt), C.c_double(gm),
                             _arr(dx), C.c_int(ro.shape[0]))

    def roe_m3_2d(self, ro, pr, vx, vy, vz, bx, by, bz, az, dt, gm, dx, dy):
        ix, jx = ro.shape
        self.lib.or2d_roe_m3(*map(_arr, (ro, pr, vx, vy, vz, bx, by, bz, az)), C.c_double(dt), C.c_double(gm),
                             _arr(dx), C.c_int(ix), _arr(dy), C.c_int(jx))

    def mlw_m_c_2d(self, ro, pr, vx, vz, bx, bz, ay, dt, qav, gm, x, xm, dx, dxm, dz, dzm):
        ix, jx = ro.shape
        self.lib.or2d_mlw_m_c(*map(_arr, (ro, pr, vx, vz, bx, bz, ay)), C.c_double(dt), C.c_double(qav),
                              C.c_double(gm), _arr(x), _arr(xm), _arr(dx), _arr(dxm), C.c_int(ix), _arr(dz), _arr(dzm),
                              C.c_int(jx))

    def mlw_m_s_2d(self, ro, pr, vx, vy, bx, by, az, dt, qav, gm, x, xm, dx, dxm, y, ym, dy, dym):
        ix, jx = ro.shape
        self.lib.or2d_mlw_m_s(*map(_arr, (ro, pr, vx, vy, bx, by, az)), C.c_double(dt), C.c_double(qav),
                              C.c_double(gm), _arr(x), _arr(xm), _arr(dx), _arr(dxm), C.c_int(ix), _arr(y), _arr(ym),
                              _arr(dy), _arr(dym), C.c_int(jx))

    def cfl_m_s_2d(self, safety, dtmin, gm, ro, pr, vx, vy, bx, by, x, dx, dy):
        ix, jx = ro.shape
        dt, merr, im, jm = C.c_double(), C.c_int(), C.c_int(), C.c_int()
        self.lib.or2d_cfl_m_s(C.byref(dt), C.c_double(safety), C.c_double(dtmin), C.byref(merr), C.c_double(gm),
                              *map(_arr, (ro, pr, vx, vy, bx, by)), _arr(x), _arr(dx), C.c_int(ix), _arr(dy),
                              C.c_int(jx), C.byref(im), C.byref(jm))
        return dt.value, merr.value, (im.value, jm.value)

    def mlw_m3_2d(self, ro, pr, vx, vy, vz, bx, by, bz, az, dt, qav, gm, dx, dxm, dy, dym):
        ix, jx = ro.shape
        self.lib.or2d_mlw_m3(*map(_arr, (ro, pr, vx, vy, vz, bx, by, bz, az)), C.c_double(dt), C.c_double(qav),
                             C.c_double(gm), _arr(dx), _arr(dxm), C.c_int(ix), _arr(dy), _arr(dym), C.c_int(jx))

    def cfl_m3_2d(self, safety, dtmin, gm, ro, pr, vx, vy, vz, bx, by, bz, dx, dy):
        ix, jx = ro.shape
        dt, merr, im, jm = C.c_double(), C.c_int(), C.c_int(), C.c_int()
        self.lib.or2d_cfl_m3(C.byref(dt), C.c_double(safety), C.c_double(dtmin), C.byref(merr), C.c_double(gm),
                             *map(_arr, (ro, pr, vx, vy, vz, bx, by, bz)), _arr(dx), C.c_int(ix), _arr(dy),
                             C.c_int(jx), C.byref(im), C.byref(jm))
        return dt.value, merr.value, (im.value, jm.value)

    # ---- UTRIP spicule model (cans1d/hdmlw/mlw_m_bg.f, md_spicule/hdmlw)
    GEOM = ("bx", "bxm", "gx", "gxm", "sc", "dsc", "scm", "dscm", "rr", "rrm", "drr", "drrm", "dx", "dxm")

    def mlw_m_bg_1d(self, ro, pr, vx, vy, by, g, dt, qav, gm):
        """g: dict of the geometry arrays bx, bxm, gx, gxm, sc, dsc, scm,
        dscm, rr, rrm, drr, drrm, dx, dxm."""
        self.lib.or1d_mlw_m_bg(*map(_arr, (ro, pr, vx, vy, by, g["bx"], g["bxm"])), C.c_double(dt),
                               C.c_double(qav), C.c_double(gm),
                               *[_arr(g[k]) for k in ("gx", "gxm", "sc", "dsc", "scm", "dscm", "rr", "rrm", "drr",
                                                      "drrm", "dx", "dxm")], C.c_int(ro.shape[0]))

    def scrdy_1d(self, dsc, dscm, sc, scm, dx, dxm):
        self.lib.or1d_scrdy(*map(_arr, (dsc, dscm, sc, scm, dx, dxm)), C.c_int(sc.shape[0]))

    def chkdav_1d(self, da, vx, floor):
        n = C.c_int()
        self.lib.or1d_chkdav(C.byref(n), _arr(da), _arr(vx), C.c_double(floor), C.c_int(da.shape[0]))
        return n.value

    def rangen(self, seed):
        """(value, new seed)"""
        s = C.c_int(int(seed))
        self.lib.or_rangen.restype = C.c_double
        v = self.lib.or_rangen(C.byref(s))
        return v, s.value

    def trqalfven(self, vy, dt, t, x, rr, seed):
        """returns the advanced seed"""
        s = C.c_int(int(seed))
        self.lib.orsp_trqalfven(_arr(vy), C.c_double(dt), C.c_double(t), _arr(x), C.c_int(vy.shape[0]), _arr(rr),
                                C.byref(s))
        return s.value

    def spicule_bnd(self, margin, ro, pr, vx, vy, by):
        self.lib.orsp_bnd(C.c_int(margin), *map(_arr, (ro, pr, vx, vy, by)), C.c_int(ro.shape[0]))

    def ref_loop_1d(self, ref, st, bx, bxm, dx, dxm, margin, gm, safety, qav, dtmin, tend, nstop):
        """The cans1d main.f loop (cfl_m; t=t+dt; mlw_m; bdfrex at both ends)
        on the state dict st (in place), calling the routines of the
        reference build `ref` (a Ref("1d")); returns (t, ns, merr)."""
        fn = lambda name: C.cast(getattr(ref.lib, name + "_"), C.c_void_p)
        t, merr = C.c_double(), C.c_int()
        ix = st["ro"].shape[0]
        ns = self.lib.orref_loop_1d(fn("cfl_m"), fn("mlw_m"), fn("bdfrex"),
                                    *map(_arr, (st["ro"], st["pr"], st["vx"], st["vy"], st["by"], bx, bxm, dx, dxm)),
                                    C.c_int(ix), C.c_int(margin), C.c_double(gm), C.c_double(safety),
                                    C.c_double(qav), C.c_double(dtmin), C.c_double(tend), C.c_int(nstop),
                                    C.byref(t), C.byref(merr))
        return t.value, ns, merr.value

    def spicule_run(self, st, g, x, margin, gm, tend, nstop, t, ns, seed, qav=1.0, safety=0.8, dtmin=1e-10,
                    floor=1e-9):
        """main.f's loop from (t, ns, seed) on the state dict st (in place);
        returns (t, ns, merr, seed)."""
        geo = (C.c_void_p * 15)(*[C.c_void_p(g[k].ctypes.data) for k in self.GEOM] + [C.c_void_p(x.ctypes.data)])
        for k in self.GEOM:
            _arr(g[k])
        _arr(x)
        tc, nc, mc, sc = C.c_double(t), C.c_int(ns), C.c_int(0), C.c_int(int(seed))
        self.lib.orsp_run(*[_arr(st[k]) for k in ("ro", "pr", "vx", "vy", "by")], geo, C.c_int(st["ro"].shape[0]),
                          C.c_int(margin), C.c_double(gm), C.c_double(qav), C.c_double(safety), C.c_double(dtmin),
                          C.c_double(floor), C.c_double(tend), C.c_int(nstop), C.byref(tc), C.byref(nc),
                          C.byref(mc), C.byref(sc))
        return tc.value, nc.value, mc.value, sc.value

    # ---- 2D
    def mlw_m_2d(self, ro, pr, vx, vy, bx, by, az, dt, qav, gm, dx, dxm, dy, dym):
        ix, jx = ro.shape
        self.lib.or2d_mlw_m(*map(_arr, (ro, pr, vx, vy, bx, by, az)), C.c_double(dt), C.c_double(qav),
                            C.c_double(gm), _arr(dx), _arr(dxm), C.c_int(ix), _arr(dy), _arr(dym), C.c_int(jx))

    def roe_m_2d(self, ro, pr, vx, vy, bx, by, az, dt, gm, dx, dy):
        ix, jx = ro.shape
        self.lib.or2d_roe_m(*map(_arr, (ro, pr, vx, vy, bx, by, az)), C.c_double(dt), C.c_double(gm),
                            _arr(dx), C.c_int(ix), _arr(dy), C.c_int(jx))

    def cfl_m_2d(self, safety, dtmin, gm, ro, pr, vx, vy, bx, by, dx, dy):
        ix, jx = ro.shape
        dt, merr, im, jm = C.c_double(), C.c_int(), C.c_int(), C.c_int()
        self.lib.or2d_cfl_m(C.byref(dt), C.c_double(safety), C.c_double(dtmin), C.byref(merr), C.c_double(gm),
                            *map(_arr, (ro, pr, vx, vy, bx, by)), _arr(dx), C.c_int(ix), _arr(dy), C.c_int(jx),
                            C.byref(im), C.byref(jm))
        return dt.value, merr.value, (im.value, jm.value)

    def bnd_2d(self, dir, kind, mbnd, margin, qq, q0=0.0, dm=None):
        ix, jx = qq.shape
        self.lib.or2d_bd(C.c_int(dir), C.c_int(kind), C.c_int(mbnd), C.c_int(margin), _arr(qq), C.c_double(q0),
                         _arr(dm), C.c_int(ix), C.c_int(jx))

    def bbtoaa_2d(self, bx, by, dxm, dym, margin):
        ix, jx = bx.shape
        az = np.zeros((ix, jx), order="F")
        self.lib.or2d_bbtoaa(_arr(az), _arr(bx), _arr(by), _arr(dxm), _arr(dym), C.c_int(margin), C.c_int(ix),
                             C.c_int(jx))
        return az

    # ---- 3D
    def mlw_m_3d(self, ro, pr, vx, vy, vz, bx, by, bz, dt, qav, gm, dx, dxm, dy, dym, dz, dzm):
        ix, jx, kx = ro.shape
        self.lib.or3d_mlw_m(*map(_arr, (ro, pr, vx, vy, vz, bx, by, bz)), C.c_double(dt), C.c_double(qav),
                            C.c_double(gm), _arr(dx), _arr(dxm), C.c_int(ix), _arr(dy), _arr(dym), C.c_int(jx),
                            _arr(dz), _arr(dzm), C.c_int(kx))

    def mlw_m_c_3d(self, ro, pr, vx, vy, vz, bx, by, bz, dt, qav, gm, x, xm, dx, dxm, dy, dym, dz, dzm):
        ix, jx, kx = ro.shape
        self.lib.or3d_mlw_m_c(*map(_arr, (ro, pr, vx, vy, vz, bx, by, bz)), C.c_double(dt), C.c_double(qav),
                              C.c_double(gm), _arr(x), _arr(xm), _arr(dx), _arr(dxm), C.c_int(ix), _arr(dy),
                              _arr(dym), C.c_int(jx), _arr(dz), _arr(dzm), C.c_int(kx))

    def mlw_m_s_3d(self, ro, pr, vx, vy, vz, bx, by, bz, dt, qav, gm, x, xm, dx, dxm, y, ym, dy, dym, dz, dzm):
        ix, jx, kx = ro.shape
        self.lib.or3d_mlw_m_s(*map(_arr, (ro, pr, vx, vy, vz, bx, by, bz)), C.c_double(dt), C.c_double(qav),
                              C.c_double(gm), _arr(x), _arr(xm), _arr(dx), _arr(dxm), C.c_int(ix), _arr(y),
                              _arr(ym), _arr(dy), _arr(dym), C.c_int(jx), _arr(dz), _arr(dzm), C.c_int(kx))

    def roe_m_3d(self, ro, pr, vx, vy, vz, bx, by, bz, dt, gm, dx, dy, dz):
        ix, jx, kx = ro.shape
        self.lib.or3d_roe_m(*map(_arr, (ro, pr, vx, vy, vz, bx, by, bz)), C.c_double(dt), C.c_double(gm),
                            _arr(dx), C.c_int(ix), _arr(dy), C.c_int(jx), _arr(dz), C.c_int(kx))

    def cfl_m_3d(self, safety, dtmin, gm, ro, pr, vx, vy, vz, bx, by, bz, dx, dy, dz):
        ix, jx, kx = ro.shape
        dt, merr, im, jm, km = C.c_double(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self.lib.or3d_cfl_m(C.byref(dt), C.c_double(safety), C.c_double(dtmin), C.byref(merr), C.c_double(gm),
                            *map(_arr, (ro, pr, vx, vy, vz, bx, by, bz)), _arr(dx), C.c_int(ix), _arr(dy),
                            C.c_int(jx), _arr(dz), C.c_int(kx), C.byref(im), C.byref(jm), C.byref(km))
        return dt.value, merr.value, (im.value, jm.value, km.value)

    def cfl_m_s_3d(self, safety, dtmin, gm, ro, pr, vx, vy, vz, bx, by, bz, x, dx, y, dy, dz):
        ix, jx, kx = ro.shape
        dt, merr, im, jm, km = C.c_double(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self.lib.or3d_cfl_m_s(C.byref(dt), C.c_double(safety), C.c_double(dtmin), C.byref(merr), C.c_double(gm),
                              *map(_arr, (ro, pr, vx, vy, vz, bx, by, bz)), _arr(x), _arr(dx), C.c_int(ix), _arr(y),
                              _arr(dy), C.c_int(jx), _arr(dz), C.c_int(kx), C.byref(im), C.byref(jm), C.byref(km))
        return dt.value, merr.value, (im.value, jm.value, km.value)

    def bnd_3d(self, dir, kind, mbnd, margin, qq, q0=0.0, dm=None):
        ix, jx, kx = qq.shape
        self.lib.or3d_bd(C.c_int(dir), C.c_int(kind), C.c_int(mbnd), C.c_int(margin), _arr(qq), C.c_double(q0),
                         _arr(dm), C.c_int(ix), C.c_int(jx), C.c_int(kx))

    def hlld_engine(self, ro, pr, vx, vy, vz, bx, by, bz, gm, mu, dt, dx, dy, dz, mfdim=(1, 1, 1)):
        ix, jx, kx = ro.shape
        mf = (C.c_int * 3)(*mfdim)
        self.lib.orh_engine(*map(_arr, (ro, pr, vx, vy, vz, bx, by, bz)), C.c_double(gm), C.c_double(mu),
                            C.c_double(dt), _arr(dx), _arr(dy), _arr(dz), C.c_int(ix), C.c_int(jx), C.c_int(kx), mf)


def _r(v, ctype=C.c_double):
    return C.byref(ctype(v))


class Ref:
    """The reference Fortran routines, compiled from source into _ref/."""

    def __init__(self, dim: str):
        path = REF_LIBS[dim]
        if not path.exists():
            raise FileNotFoundError(f"{path} not built (make -C oracle ref)")
        self.lib = C.CDLL(str(path))
        self.dim = dim

    def call(self, name, *args):
        """Call a reference routine by its Fortran name; numpy arrays go by
        address, Python floats as REAL*8, ints as INTEGER*4 (by reference)."""
        cargs = []
        keep = []
        for a in args:
            if isinstance(a, np.ndarray):
                if a.dtype == np.int32:
                    cargs.append(C.c_void_p(a.ctypes.data))
                else:
                    cargs.append(_arr(a))
            elif isinstance(a, (C.c_double, C.c_int)):
                cargs.append(C.byref(a))
                keep.append(a)
            elif isinstance(a, bool) or isinstance(a, (int, np.integer)):
                v = C.c_int(int(a))
                keep.append(v)
                cargs.append(C.byref(v))
            else:
                v = C.c_double(float(a))
                keep.append(v)
                cargs.append(C.byref(v))
        fn = getattr(self.lib, name + "_")
        big_stack_call(fn, *cargs)
