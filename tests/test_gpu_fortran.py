"""The Fortran drop-in: CANS main programs (utrip_amd/fortran/cans{1,2}d_main.f90)
that call mlw_m / roe_m / cfl_m / bd* with the reference's argument lists,
linked against libcans{1,2}d_gpu.so (ISO_C_BINDING over the C-ABI), restart
from DAC files and write DAC files.  Their output must equal the C
restatement's main loop bit for bit, in the strict (copy every call) and the
resident (arrays stay on the GPU) modes."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from tests.helpers import assert_bitwise
from utrip_amd import PKG_DIR, dac
from utrip_amd import problems as P

pytestmark = pytest.mark.gpu

BIN = PKG_DIR / "bin"
F1 = ("ro", "pr", "vx", "vy", "by")
F2 = ("ro", "pr", "vx", "vy", "bx", "by", "az")


def run_main(args, timeout=120):
    print("CASE fortran", " ".join(map(str, args)), flush=True)
    r = subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def oracle_loop_1d(oracle, pb, engine, nsteps, margin):
    f = {k: v.copy() for k, v in pb.fields.items()}
    g, pp = pb.grid, pb.params
    t = 0.0
    for _ in range(nsteps):
        dt, merr, _ = oracle.cfl_m_1d(pp["safety"], 1e-10, pp["gm"], f["bx"], *[f[k] for k in ("ro", "pr", "vx", "vy", "by")],
                                      g["dx"])
        t = t + dt
        if engine == "mlw":
            oracle.mlw_m_1d(*[f[k] for k in F1], f["bx"], f["bxm"], dt, pp["qav"], pp["gm"], g["dx"], g["dxm"])
        else:
            oracle.roe_m_1d(*[f[k] for k in F1], f["bx"], f["bxm"], dt, pp["gm"], g["dx"])
        for side in (0, 1):
            for k in F1:
                oracle.bnd_1d(0, side, margin, f[k])
    return f, t


@pytest.mark.parametrize("engine,resident", [("mlw", 0), ("mlw", 1), ("roe", 1)])
def test_cans1d_main(dev, oracle, tmp_path, engine, resident):
    margin = 2 if engine == "mlw" else 3
    pb = P.shock_tube_1d(nx=200, margin=margin)
    nsteps = 50
    ref, t = oracle_loop_1d(oracle, pb, engine, nsteps, margin)
    ind, outd = tmp_path / "in", tmp_path / "out"
    dac.write_state(ind, {k: pb.fields[k] for k in F1 + ("bx", "bxm")}, {"dx": pb.grid["dx"], "dxm": pb.grid["dxm"]})
    outd.mkdir()
    pp = pb.params
    out = run_main([BIN / "cans1d_main", engine, "fre", ind, outd, nsteps, 1e30, margin, repr(pp["safety"]),
                    repr(pp["qav"]), repr(pp["gm"]), resident])
    assert f"step={nsteps:8d}" in out, out
    for k in F1:
        assert_bitwise(dac.read_last(outd / f"{k}.dac"), ref[k], k)
    assert float(dac.read_last(outd / "t.dac")) == t


@pytest.mark.parametrize("engine,resident", [("mlw", 0), ("mlw", 1), ("roe", 1)])
def test_cans2d_main(dev, oracle, tmp_path, engine, resident):
    margin = 1 if engine == "mlw" else 3
    pb = P.orszag_tang_2d(n=40, margin=margin)
    nsteps = 12
    f = {k: v.copy(order="F") for k, v in pb.fields.items()}
    g, pp = pb.grid, pb.params
    t = 0.0
    for _ in range(nsteps):
        dt, merr, _ = oracle.cfl_m_2d(pp["safety"], 1e-10, pp["gm"], *[f[k] for k in F2[:6]], g["dx"], g["dy"])
        t = t + dt
        if engine == "mlw":
            oracle.mlw_m_2d(*[f[k] for k in F2], dt, pp["qav"], pp["gm"], g["dx"], g["dxm"], g["dy"], g["dym"])
        else:
            oracle.roe_m_2d(*[f[k] for k in F2], dt, pp["gm"], g["dx"], g["dy"])
        for d in (0, 1):
            for k in F2:
                oracle.bnd_2d(d, 7, margin, margin, f[k])
    ind, outd = tmp_path / "in", tmp_path / "out"
    dac.write_state(ind, pb.fields, {k: g[k] for k in ("dx", "dxm", "dy", "dym")})
    outd.mkdir()
    out = run_main([BIN / "cans2d_main", engine, ind, outd, nsteps, 1e30, margin, repr(pp["safety"]),
                    repr(pp["qav"]), repr(pp["gm"]), resident])
    assert f"step={nsteps:8d}" in out, out
    for k in F2:
        assert_bitwise(dac.read_last(outd / f"{k}.dac"), f[k], k)
    assert float(dac.read_last(outd / "t.dac")) == t


@pytest.mark.parametrize("resident", [0, 1])
def test_spicule_main(dev, oracle, tmp_path, resident):
    """The UTRIP spicule loop as a Fortran main program on the drop-in
    library (mlw_m_bg, trqalfven, chkdav, cfl_m, bd*), against the C
    restatement's loop."""
    from utrip_amd import spicule
    sp = spicule.setup()
    nsteps, seed = 60, 246810
    st = {k: np.asfortranarray(v.copy()) for k, v in sp.state.items()}
    g = {k: np.asfortranarray(v) for k, v in sp.geom.items()}
    t, ns, merr, _ = oracle.spicule_run(st, g, np.asfortranarray(sp.x), sp.margin, sp.gm, spicule.TEND, nsteps,
                                        0.0, 0, seed)
    assert (ns, merr) == (nsteps, 0)
    ind, outd = tmp_path / "in", tmp_path / "out"
    dac.write_state(ind, sp.state, dict(sp.geom, x=sp.x))
    outd.mkdir()
    out = run_main([BIN / "spicule_main", ind, outd, nsteps, repr(spicule.TEND), repr(sp.gm), seed, resident])
    assert f"step={nsteps:8d}" in out, out
    for k in F1:
        assert_bitwise(dac.read_last(outd / f"{k}.dac"), st[k], k)
    assert float(dac.read_last(outd / "t.dac")) == t


def test_dropin_mlw_m3_by_fortran_names(dev, oracle):
    """mlw_m3_, roe_m3_, cfl_m3_, mlw_m_c_, mlw_m_s_ and cfl_m_s_ of libcans{1,2,3}d_gpu.so called as a Fortran main
    program calls them (every argument by reference, host arrays, strict
    mode), against the restatement."""
    import ctypes as C

    from tests.helpers import stretched_grid
    from tests.test_oracle_m3 import state1, state2

    def call(lib, name, *args):
        keep, cargs = [], []
        for a in args:
            if isinstance(a, np.ndarray):
                cargs.append(C.c_void_p(a.ctypes.data))
            elif isinstance(a, (C.c_double, C.c_int)):
                cargs.append(C.byref(a))
            elif isinstance(a, int):
                v = C.c_int(a); keep.append(v); cargs.append(C.byref(v))
            else:
                v = C.c_double(a); keep.append(v); cargs.append(C.byref(v))
        getattr(lib, name + "_")(*cargs)

    lib2 = C.CDLL(str(PKG_DIR / "libcans2d_gpu.so"))
    ix, jx = 70, 41
    s = state2(ix, jx, 4)
    _, dx, dxm = stretched_grid(ix, seed=1)
    _, dy, dym = stretched_grid(jx, seed=2)
    F = ("ro", "pr", "vx", "vy", "vz", "bx", "by", "bz", "az")
    a = {k: v.copy(order="F") for k, v in s.items()}
    b = {k: v.copy(order="F") for k, v in s.items()}
    oracle.mlw_m3_2d(*[a[k] for k in F], 1.1e-3, 3.0, 5 / 3, dx, dxm, dy, dym)
    call(lib2, "mlw_m3", *[b[k] for k in F], 1.1e-3, 3.0, 5 / 3, dx, dxm, ix, dy, dym, jx)
    for k in F:
        assert_bitwise(b[k], a[k], k)
    dt, merr = C.c_double(), C.c_int()
    call(lib2, "cfl_m3", dt, 0.4, 1e-10, merr, 5 / 3, *[s[k] for k in F[:8]], dx, ix, dy, jx)
    want = oracle.cfl_m3_2d(0.4, 1e-10, 5 / 3, *[s[k] for k in F[:8]], dx, dy)
    assert (dt.value, merr.value) == want[:2]
    a = {k: v.copy(order="F") for k, v in s.items()}
    b = {k: v.copy(order="F") for k, v in s.items()}
    oracle.roe_m3_2d(*[a[k] for k in F], 9e-4, 5 / 3, dx, dy)
    call(lib2, "roe_m3", *[b[k] for k in F], 9e-4, 5 / 3, dx, ix, dy, jx)
    for k in F:   # Roe: the interior 4..n-3 (see tests/test_gpu_m3.py)
        assert_bitwise(b[k][3:-3, 3:-3], a[k][3:-3, 3:-3], k)
    from tests.test_oracle_m3 import FC, state_c
    sc, x, xm, dxc, dxmc = state_c(ix, jx, 6)
    a = {k: v.copy(order="F") for k, v in sc.items()}
    b = {k: v.copy(order="F") for k, v in sc.items()}
    oracle.mlw_m_c_2d(*[a[k] for k in FC], 1e-3, 3.0, 5 / 3, x, xm, dxc, dxmc, dy, dym)
    call(lib2, "mlw_m_c", *[b[k] for k in FC], 1e-3, 3.0, 5 / 3, x, xm, dxc, dxmc, ix, dy, dym, jx)
    for k in FC:
        assert_bitwise(b[k], a[k], k)
    from tests.test_oracle_m3 import F2S, state_s2
    ss2, xs, xms, dxs, dxms, ys, yms, dys, dyms = state_s2(ix, jx, 8)
    a = {k: v.copy(order="F") for k, v in ss2.items()}
    b = {k: v.copy(order="F") for k, v in ss2.items()}
    oracle.mlw_m_s_2d(*[a[k] for k in F2S], 1e-3, 3.0, 5 / 3, xs, xms, dxs, dxms, ys, yms, dys, dyms)
    call(lib2, "mlw_m_s", *[b[k] for k in F2S], 1e-3, 3.0, 5 / 3, xs, xms, dxs, dxms, ix, ys, yms, dys, dyms, jx)
    for k in F2S:
        assert_bitwise(b[k], a[k], k)
    f6 = ("ro", "pr", "vx", "vy", "bx", "by")
    ref = oracle.cfl_m_s_2d(0.4, 1e-10, 5 / 3, *[ss2[k] for k in f6], xs, dxs, dys)
    dt, merr = C.c_double(), C.c_int()
    call(lib2, "cfl_m_s", dt, 0.4, 1e-10, merr, 5 / 3, *[ss2[k] for k in f6], xs, dxs, ix, dys, jx)
    assert dt.value == ref[0] and merr.value == ref[1]
    lib1 = C.CDLL(str(PKG_DIR / "libcans1d_gpu.so"))
    ix = 123
    s1 = state1(ix, 2)
    _, dx, dxm = stretched_grid(ix, seed=3)
    F1m = ("ro", "pr", "vx", "vy", "vz", "by", "bz")
    a = {k: v.copy() for k, v in s1.items()}
    b = {k: v.copy() for k, v in s1.items()}
    oracle.mlw_m3_1d(*[a[k] for k in F1m], a["bx"], a["bxm"], 2e-3, 3.0, 5 / 3, dx, dxm)
    call(lib1, "mlw_m3", *[b[k] for k in F1m], b["bx"], b["bxm"], 2e-3, 3.0, 5 / 3, dx, dxm, ix)
    for k in F1m:
        assert_bitwise(b[k], a[k], k)
    a = {k: v.copy() for k, v in s1.items()}
    b = {k: v.copy() for k, v in s1.items()}
    oracle.roe_m3_1d(*[a[k] for k in F1m], a["bx"], a["bxm"], 1e-3, 5 / 3, dx)
    call(lib1, "roe_m3", *[b[k] for k in F1m], b["bx"], b["bxm"], 1e-3, 5 / 3, dx, ix)
    for k in F1m:
        assert_bitwise(b[k][3:-3], a[k][3:-3], k)
    # 3D cylindrical mlw_m_c_ of libcans3d_gpu.so (md_mhdsn/hdmlw/main.f:179)
    from tests.helpers import grid3, smooth_state_3d
    from tests.test_oracle_m3 import F3, radii
    lib3 = C.CDLL(str(PKG_DIR / "libcans3d_gpu.so"))
    shape = (17, 9, 11)
    s3 = smooth_state_3d(*shape, seed=41)
    dx, dxm, dy, dym, dz, dzm = grid3(*shape)
    x, xm = radii(shape[0], 0.25, dx)
    a = {k: v.copy(order="F") for k, v in s3.items()}
    b = {k: v.copy(order="F") for k, v in s3.items()}
    oracle.mlw_m_c_3d(*[a[k] for k in F3], 1e-3, 3.0, 5 / 3, x, xm, dx, dxm, dy, dym, dz, dzm)
    call(lib3, "mlw_m_c", *[b[k] for k in F3], 1e-3, 3.0, 5 / 3, x, xm, dx, dxm, shape[0], dy, dym, shape[1],
         dz, dzm, shape[2])
    for k in F3:
        assert_bitwise(b[k], a[k], k)
    # 3D spherical mlw_m_s_ and cfl_m_s_
    from tests.test_oracle_m3 import polar
    y, ym = polar(shape[1], 0.7, dy)
    a = {k: v.copy(order="F") for k, v in s3.items()}
    b = {k: v.copy(order="F") for k, v in s3.items()}
    oracle.mlw_m_s_3d(*[a[k] for k in F3], 1e-3, 3.0, 5 / 3, x, xm, dx, dxm, y, ym, dy, dym, dz, dzm)
    call(lib3, "mlw_m_s", *[b[k] for k in F3], 1e-3, 3.0, 5 / 3, x, xm, dx, dxm, shape[0], y, ym, dy, dym,
         shape[1], dz, dzm, shape[2])
    for k in F3:
        assert_bitwise(b[k], a[k], k)
    ref = oracle.cfl_m_s_3d(0.4, 1e-10, 5 / 3, *[s3[k] for k in F3], x, dx, y, dy, dz)
    dt, merr = C.c_double(), C.c_int()
    call(lib3, "cfl_m_s", dt, 0.4, 1e-10, merr, 5 / 3, *[s3[k] for k in F3], x, dx, shape[0], y, dy, shape[1],
         dz, shape[2])
    assert dt.value == ref[0] and merr.value == ref[1]


@pytest.mark.gpu
def test_mirror_remap_smaller_array(dev):
    """The host mirror of the Fortran layer is keyed by the host address; an
    address that comes back for a smaller array (the allocator reuses it)
    must be copied back with the smaller size, never past its end."""
    import ctypes as C
    lib = C.CDLL(str(PKG_DIR / "libcans_gpu.so"))
    lib.cans_gpu_map.restype = C.c_void_p
    lib.cans_gpu_map.argtypes = [C.c_void_p, C.c_size_t, C.c_int]
    lib.cans_gpu_unmap.argtypes = [C.c_void_p, C.c_int]
    big = np.arange(4096, dtype=np.float64)
    addr = big.ctypes.data
    assert lib.cans_gpu_map(addr, big.nbytes, 1)
    assert lib.cans_gpu_unmap(addr, 1) == 0
    n = 1000
    big[n:] = -7.0                       # host memory past the smaller array
    small = big[:n]
    assert small.ctypes.data == addr
    assert lib.cans_gpu_map(addr, small.nbytes, 1)
    assert lib.cans_gpu_unmap(addr, 1) == 0
    assert np.array_equal(big[:n], np.arange(n, dtype=np.float64))
    assert np.all(big[n:] == -7.0)
