"""The three-component engines on the GPU against the C restatement:
mlw_m3 and cfl_m3 in 1D and 2D as single calls (ragged and exact tile
shapes) and as a resident loop (cans_run with CANS_ENGINE_MLW3), bit for
bit in FP64."""
import numpy as np
import pytest
import torch

from tests.helpers import assert_bitwise, stretched_grid, to_dev, to_host
from tests.test_oracle_m3 import F1, F2, state1, state2
from utrip_amd import cans

pytestmark = pytest.mark.gpu


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)


@pytest.mark.parametrize("ix", [5, 64, 301, 1030])
def test_mlw_m3_1d_call(oracle, dev, ix):
    print(f"mlw_m3 1d ix={ix}", flush=True)
    s = state1(ix, ix)
    _, dx, dxm = stretched_grid(ix, seed=ix)
    a = {k: v.copy() for k, v in s.items()}
    oracle.mlw_m3_1d(*[a[k] for k in F1], a["bx"], a["bxm"], 1.7e-3, 3.0, 5 / 3, dx, dxm)
    d = {k: T(v, dev) for k, v in s.items()}
    cans.mlw_m3_1d(*[d[k] for k in F1], d["bx"], d["bxm"], 1.7e-3, 3.0, 5 / 3, T(dx, dev), T(dxm, dev))
    torch.cuda.synchronize()
    for k in F1:
        assert_bitwise(d[k].cpu().numpy(), a[k], k)


def test_cfl_m3_1d_call(oracle, dev):
    ix = 300
    s = state1(ix, 8)
    _, dx, _ = stretched_grid(ix)
    want = oracle.cfl_m3_1d(0.4, 1e-10, 5 / 3, s["bx"], s["ro"], s["pr"], s["vx"], s["vy"], s["by"], s["vz"], s["bz"],
                            dx)
    d = {k: T(v, dev) for k, v in s.items()}
    got = cans.cfl_m3_1d(0.4, 1e-10, 5 / 3, d["bx"], d["ro"], d["pr"], d["vx"], d["vy"], d["by"], d["vz"], d["bz"],
                         T(dx, dev))
    assert got == want[:2]


@pytest.mark.parametrize("ix,jx", [(5, 6), (33, 9), (40, 33), (97, 64), (130, 71)])
def test_mlw_m3_2d_call(oracle, dev, ix, jx):
    print(f"mlw_m3 2d {ix}x{jx}", flush=True)
    s = state2(ix, jx, ix + jx)
    _, dx, dxm = stretched_grid(ix, seed=ix)
    _, dy, dym = stretched_grid(jx, seed=jx + 1)
    a = {k: v.copy(order="F") for k, v in s.items()}
    oracle.mlw_m3_2d(*[a[k] for k in F2], 1.3e-3, 3.0, 5 / 3, dx, dxm, dy, dym)
    d = {k: to_dev(v, dev) for k, v in s.items()}
    cans.mlw_m3_2d(*[d[k] for k in F2], 1.3e-3, 3.0, 5 / 3, T(dx, dev), T(dxm, dev), T(dy, dev), T(dym, dev))
    torch.cuda.synchronize()
    for k in F2:
        assert_bitwise(to_host(d[k]), a[k], k)


def test_cfl_m3_2d_call(oracle, dev):
    ix, jx = 70, 45
    s = state2(ix, jx, 12)
    _, dx, _ = stretched_grid(ix)
    _, dy, _ = stretched_grid(jx, seed=4)
    names = ("ro", "pr", "vx", "vy", "vz", "bx", "by", "bz")
    want = oracle.cfl_m3_2d(0.4, 1e-10, 5 / 3, *[s[k] for k in names], dx, dy)
    d = {k: to_dev(s[k], dev) for k in names}
    got = cans.cfl_m3_2d(0.4, 1e-10, 5 / 3, *[d[k] for k in names], T(dx, dev), T(dy, dev))
    assert got == want[:2]


def periodic2(oracle, f, margin):
    for dr in (0, 1):
        for k in F2:
            oracle.bnd_2d(dr, 7, margin, margin, f[k])


@pytest.mark.parametrize("margin,ghost_min", [(1, False), (2, True), (1, True)])
def test_mlw_m3_2d_resident_loop(oracle, dev, margin, ghost_min):
    """40 steps of {cfl_m3; t=t+dt; mlw_m3; bdper x, y} in the resident loop
    against the restatement's loop (periodic 2.5D MHD with all components).
    In the periodic box the kernel advances the clock and takes the next
    step's cfl_m3 (the first step of each call a cfl pass); with ghost_min
    the smallest dx and dy lie in the end ghost cells, so the reference's
    cfl_m3 takes dt from a ghost image.  Issued in two calls."""
    ix, jx, nsteps = 66, 40, 40
    s = state2(ix, jx, 77)
    _, dx, dxm = stretched_grid(ix, seed=5)
    _, dy, dym = stretched_grid(jx, seed=6)
    if ghost_min:
        dx[0], dx[ix - 1] = 0.3 * dx.min(), 0.35 * dx.min()
        dy[0], dy[jx - 1] = 0.35 * dy.min(), 0.3 * dy.min()
    periodic2(oracle, s, margin)
    a = {k: v.copy(order="F") for k, v in s.items()}
    t = 0.0
    names = ("ro", "pr", "vx", "vy", "vz", "bx", "by", "bz")
    for _ in range(nsteps):
        dt, merr, _ = oracle.cfl_m3_2d(0.4, 1e-10, 5 / 3, *[a[k] for k in names], dx, dy)
        assert merr == 0
        t = t + dt
        oracle.mlw_m3_2d(*[a[k] for k in F2], dt, 3.0, 5 / 3, dx, dxm, dy, dym)
        periodic2(oracle, a, margin)
    st = [to_dev(s[k], dev) for k in F2]
    grid = {k: T(v, dev) for k, v in dict(dx=dx, dxm=dxm, dy=dy, dym=dym).items()}
    run = cans.Run(2, cans.ENGINE_MLW3, st, grid, 5 / 3, 0.4, qav=3.0, margin=margin, bc=7)
    run.reset()
    run.steps(nsteps // 2)
    run.steps(nsteps - nsteps // 2)
    cur, info = run.state()
    torch.cuda.synchronize()
    assert info["nexec"] == nsteps and info["merr"] == 0 and info["t"] == t
    for k, g in zip(F2, cur):
        assert_bitwise(to_host(g), a[k], k)


def test_mlw_m3_1d_resident_loop(oracle, dev):
    ix, margin, nsteps = 200, 2, 60
    s = state1(ix, 21)
    _, dx, dxm = stretched_grid(ix, seed=9)
    for side in (0, 1):
        for k in F1:
            oracle.bnd_1d(0, side, margin, s[k])
    a = {k: v.copy() for k, v in s.items()}
    t = 0.0
    for _ in range(nsteps):
        dt, merr, _ = oracle.cfl_m3_1d(0.4, 1e-10, 5 / 3, a["bx"], a["ro"], a["pr"], a["vx"], a["vy"], a["by"], a["vz"],
                                       a["bz"], dx)
        t = t + dt
        oracle.mlw_m3_1d(*[a[k] for k in F1], a["bx"], a["bxm"], dt, 3.0, 5 / 3, dx, dxm)
        for side in (0, 1):
            for k in F1:
                oracle.bnd_1d(0, side, margin, a[k])
    st = [T(s[k], dev) for k in F1]
    grid = {k: T(v, dev) for k, v in dict(dx=dx, dxm=dxm).items()}
    run = cans.Run(1, cans.ENGINE_MLW3, st, grid, 5 / 3, 0.4, qav=3.0, margin=margin, bc=0,
                   bx1d=T(s["bx"], dev), bxm1d=T(s["bxm"], dev))
    run.reset()
    run.steps(nsteps)
    cur, info = run.state()
    torch.cuda.synchronize()
    assert info["nexec"] == nsteps and info["t"] == t
    for k, g in zip(F1, cur):
        assert_bitwise(g.cpu().numpy(), a[k], k)


@pytest.mark.parametrize("ix", [9, 80, 1030])
def test_roe_m3_1d_call(oracle, dev, ix):
    print(f"roe_m3 1d ix={ix}", flush=True)
    s = state1(ix, ix + 3)
    _, dx, _ = stretched_grid(ix, seed=ix)
    a = {k: v.copy() for k, v in s.items()}
    oracle.roe_m3_1d(*[a[k] for k in F1], a["bx"], a["bxm"], 1.1e-3, 5 / 3, dx)
    d = {k: T(v, dev) for k, v in s.items()}
    cans.roe_m3_1d(*[d[k] for k in F1], d["bx"], d["bxm"], 1.1e-3, 5 / 3, T(dx, dev))
    torch.cuda.synchronize()
    # cells 3 and ix-2 read the half step of cells 1 and ix, which roe_m3.f
    # never writes (margin >= 3 in its drivers): compare 4..ix-3
    for k in F1:
        assert_bitwise(d[k].cpu().numpy()[3:-3], a[k][3:-3], k)


@pytest.mark.parametrize("ix,jx", [(9, 10), (40, 33), (130, 71)])
def test_roe_m3_2d_call(oracle, dev, ix, jx):
    print(f"roe_m3 2d {ix}x{jx}", flush=True)
    s = state2(ix, jx, 2 * ix + jx)
    _, dx, _ = stretched_grid(ix, seed=ix)
    _, dy, _ = stretched_grid(jx, seed=jx + 2)
    a = {k: v.copy(order="F") for k, v in s.items()}
    oracle.roe_m3_2d(*[a[k] for k in F2], 9e-4, 5 / 3, dx, dy)
    d = {k: to_dev(v, dev) for k, v in s.items()}
    cans.roe_m3_2d(*[d[k] for k in F2], 9e-4, 5 / 3, T(dx, dev), T(dy, dev))
    torch.cuda.synchronize()
    for k in F2:   # as in 1D: the interior 4..n-3 in each direction
        assert_bitwise(to_host(d[k])[3:-3, 3:-3], a[k][3:-3, 3:-3], k)


def test_roe_m3_2d_resident_loop(oracle, dev):
    """30 steps of {cfl_m3; t=t+dt; roe_m3; bdper x, y} (margin 3)."""
    ix, jx, margin, nsteps = 50, 38, 3, 30
    s = state2(ix, jx, 61)
    _, dx, _ = stretched_grid(ix, seed=5)
    _, dy, _ = stretched_grid(jx, seed=6)
    periodic2(oracle, s, margin)
    a = {k: v.copy(order="F") for k, v in s.items()}
    t = 0.0
    names = ("ro", "pr", "vx", "vy", "vz", "bx", "by", "bz")
    for _ in range(nsteps):
        dt, merr, _ = oracle.cfl_m3_2d(0.4, 1e-10, 5 / 3, *[a[k] for k in names], dx, dy)
        t = t + dt
        oracle.roe_m3_2d(*[a[k] for k in F2], dt, 5 / 3, dx, dy)
        periodic2(oracle, a, margin)
    st = [to_dev(s[k], dev) for k in F2]
    grid = {"dx": T(dx, dev), "dy": T(dy, dev)}
    run = cans.Run(2, cans.ENGINE_ROE3, st, grid, 5 / 3, 0.4, margin=margin, bc=7)
    run.reset()
    run.steps(nsteps)
    cur, info = run.state()
    torch.cuda.synchronize()
    assert info["nexec"] == nsteps and info["t"] == t
    for k, g in zip(F2, cur):
        assert_bitwise(to_host(g), a[k], k)


def test_roe_m3_1d_resident_loop(oracle, dev):
    ix, margin, nsteps = 160, 3, 40
    s = state1(ix, 33)
    _, dx, dxm = stretched_grid(ix, seed=10)
    for side in (0, 1):
        for k in F1:
            oracle.bnd_1d(0, side, margin, s[k])
    a = {k: v.copy() for k, v in s.items()}
    t = 0.0
    for _ in range(nsteps):
        dt, merr, _ = oracle.cfl_m3_1d(0.4, 1e-10, 5 / 3, a["bx"], a["ro"], a["pr"], a["vx"], a["vy"], a["by"], a["vz"],
                                       a["bz"], dx)
        t = t + dt
        oracle.roe_m3_1d(*[a[k] for k in F1], a["bx"], a["bxm"], dt, 5 / 3, dx)
        for side in (0, 1):
            for k in F1:
                oracle.bnd_1d(0, side, margin, a[k])
    st = [T(s[k], dev) for k in F1]
    run = cans.Run(1, cans.ENGINE_ROE3, st, {"dx": T(dx, dev), "dxm": T(dxm, dev)}, 5 / 3, 0.4, margin=margin, bc=0,
                   bx1d=T(s["bx"], dev), bxm1d=T(s["bxm"], dev))
    run.reset()
    run.steps(nsteps)
    cur, info = run.state()
    torch.cuda.synchronize()
    assert info["nexec"] == nsteps and info["t"] == t
    for k, g in zip(F1, cur):
        assert_bitwise(g.cpu().numpy(), a[k], k)


from tests.test_oracle_m3 import FC, state_c  # noqa: E402


@pytest.mark.parametrize("ix,jx", [(5, 6), (33, 9), (70, 51), (130, 71)])
def test_mlw_m_c_call(oracle, dev, ix, jx):
    print(f"mlw_m_c {ix}x{jx}", flush=True)
    s, x, xm, dx, dxm = state_c(ix, jx, ix + 3 * jx)
    _, dz, dzm = stretched_grid(jx, seed=jx + 5)
    a = {k: v.copy(order="F") for k, v in s.items()}
    oracle.mlw_m_c_2d(*[a[k] for k in FC], 1.2e-3, 3.0, 5 / 3, x, xm, dx, dxm, dz, dzm)
    d = {k: to_dev(v, dev) for k, v in s.items()}
    cans.mlw_m_c_2d(*[d[k] for k in FC], 1.2e-3, 3.0, 5 / 3, *[T(v, dev) for v in (x, xm, dx, dxm, dz, dzm)])
    torch.cuda.synchronize()
    for k in FC:
        assert_bitwise(to_host(d[k]), a[k], k)


def test_mlw_m_c_axis_face_call(oracle, dev):
    ix, jx = 24, 20
    s, x, xm, dx, dxm = state_c(ix, jx, 3, r0=0.0)
    x = np.asfortranarray(x - x[1] + 0.5 * dxm[1])
    xm = np.asfortranarray(np.append(0.5 * (x[:-1] + x[1:]), x[-1] + 0.5 * dx[-1]))
    xm[0] = 0.0
    _, dz, dzm = stretched_grid(jx, seed=8)
    a = {k: v.copy(order="F") for k, v in s.items()}
    oracle.mlw_m_c_2d(*[a[k] for k in FC], 1e-3, 3.0, 5 / 3, x, xm, dx, dxm, dz, dzm)
    d = {k: to_dev(v, dev) for k, v in s.items()}
    cans.mlw_m_c_2d(*[d[k] for k in FC], 1e-3, 3.0, 5 / 3, *[T(v, dev) for v in (x, xm, dx, dxm, dz, dzm)])
    torch.cuda.synchronize()
    for k in FC:
        assert_bitwise(to_host(d[k]), a[k], k)


def test_mlw_m_c_resident_loop(oracle, dev):
    """30 steps of {cfl_m (vx, vz, bx, bz); t=t+dt; mlw_m_c; bdper x, z}."""
    ix, jx, margin, nsteps = 50, 40, 1, 30
    s, x, xm, dx, dxm = state_c(ix, jx, 17)
    _, dz, dzm = stretched_grid(jx, seed=18)
    for dr in (0, 1):
        for k in FC:
            oracle.bnd_2d(dr, 7, margin, margin, s[k])
    a = {k: v.copy(order="F") for k, v in s.items()}
    t = 0.0
    for _ in range(nsteps):
        dt, merr, _ = oracle.cfl_m_2d(0.4, 1e-10, 5 / 3, a["ro"], a["pr"], a["vx"], a["vz"], a["bx"], a["bz"], dx, dz)
        t = t + dt
        oracle.mlw_m_c_2d(*[a[k] for k in FC], dt, 3.0, 5 / 3, x, xm, dx, dxm, dz, dzm)
        for dr in (0, 1):
            for k in FC:
                oracle.bnd_2d(dr, 7, margin, margin, a[k])
    st = [to_dev(s[k], dev) for k in FC]
    grid = {"dx": T(dx, dev), "dxm": T(dxm, dev), "dy": T(dz, dev), "dym": T(dzm, dev)}
    run = cans.Run(2, cans.ENGINE_MLWC, st, grid, 5 / 3, 0.4, qav=3.0, margin=margin, bc=7, bx1d=T(x, dev),
                   bxm1d=T(xm, dev))
    run.reset()
    run.steps(nsteps)
    cur, info = run.state()
    torch.cuda.synchronize()
    assert info["nexec"] == nsteps and info["t"] == t
    for k, g in zip(FC, cur):
        assert_bitwise(to_host(g), a[k], k)


# ---------------------------------------------------------------- mlw_m_c, 3D (r, phi, z)
from tests.helpers import grid3, smooth_state_3d  # noqa: E402
from tests.test_oracle_m3 import F3, radii  # noqa: E402


@pytest.mark.parametrize("shape", [(3, 3, 3), (6, 5, 7), (21, 8, 12), (40, 33, 20)])
def test_mlw_m_c_3d_call(oracle, dev, shape):
    s = smooth_state_3d(*shape, seed=sum(shape) + 1)
    dx, dxm, dy, dym, dz, dzm = grid3(*shape)
    x, xm = radii(shape[0], 0.3, dx)
    a = {k: v.copy(order="F") for k, v in s.items()}
    oracle.mlw_m_c_3d(*[a[k] for k in F3], 1e-3, 3.0, 5 / 3, x, xm, dx, dxm, dy, dym, dz, dzm)
    d = {k: to_dev(v, dev) for k, v in s.items()}
    cans.mlw_m_c_3d(*[d[k] for k in F3], 1e-3, 3.0, 5 / 3, *[T(v, dev) for v in (x, xm, dx, dxm, dy, dym, dz, dzm)])
    torch.cuda.synchronize()
    for k in F3:
        assert_bitwise(to_host(d[k]), a[k], k)


def test_mlw_m_c_3d_axis_face_call(oracle, dev):
    shape = (10, 7, 6)
    s = smooth_state_3d(*shape, seed=17)
    dx, dxm, dy, dym, dz, dzm = grid3(*shape)
    x, xm = radii(shape[0], 0.0, dx)
    xm[0] = 0.0
    a = {k: v.copy(order="F") for k, v in s.items()}
    oracle.mlw_m_c_3d(*[a[k] for k in F3], 1e-3, 3.0, 5 / 3, x, xm, dx, dxm, dy, dym, dz, dzm)
    d = {k: to_dev(v, dev) for k, v in s.items()}
    cans.mlw_m_c_3d(*[d[k] for k in F3], 1e-3, 3.0, 5 / 3, *[T(v, dev) for v in (x, xm, dx, dxm, dy, dym, dz, dzm)])
    torch.cuda.synchronize()
    for k in F3:
        assert_bitwise(to_host(d[k]), a[k], k)
    assert not np.isfinite(a["ro"][1, 1:-1, 1:-1]).all()


def test_mlw_m_c_3d_resident_loop(oracle, dev):
    """12 steps of {cfl_m (3D); t=t+dt; mlw_m_c; bdper x, y, z} -- the
    md_mhdsn 3D loop body (main.f:169-183) with periodic faces."""
    shape, margin, nsteps = (24, 20, 16), 1, 12
    s = smooth_state_3d(*shape, seed=31)
    dx, dxm, dy, dym, dz, dzm = grid3(*shape)
    x, xm = radii(shape[0], 0.5, dx)
    for dr in (0, 1, 2):
        for k in F3:
            oracle.bnd_3d(dr, 7, margin, margin, s[k])
    a = {k: v.copy(order="F") for k, v in s.items()}
    t = 0.0
    for _ in range(nsteps):
        dt, merr, _ = oracle.cfl_m_3d(0.4, 1e-10, 5 / 3, *[a[k] for k in F3], dx, dy, dz)
        t = t + dt
        oracle.mlw_m_c_3d(*[a[k] for k in F3], dt, 3.0, 5 / 3, x, xm, dx, dxm, dy, dym, dz, dzm)
        for dr in (0, 1, 2):
            for k in F3:
                oracle.bnd_3d(dr, 7, margin, margin, a[k])
    st = [to_dev(s[k], dev) for k in F3]
    grid = {k: T(v, dev) for k, v in zip(("dx", "dxm", "dy", "dym", "dz", "dzm"), (dx, dxm, dy, dym, dz, dzm))}
    run = cans.Run(3, cans.ENGINE_MLWC, st, grid, 5 / 3, 0.4, qav=3.0, margin=margin, bc=7, bx1d=T(x, dev),
                   bxm1d=T(xm, dev))
    run.reset()
    run.steps(nsteps)
    cur, info = run.state()
    torch.cuda.synchronize()
    assert info["nexec"] == nsteps and info["t"] == t
    for k, g in zip(F3, cur):
        assert_bitwise(to_host(g), a[k], k)


# ---------------------------------------------------------------- mlw_m_s, cfl_m_s, 3D (r, theta, phi)
from tests.test_oracle_m3 import polar  # noqa: E402


@pytest.mark.parametrize("shape", [(3, 3, 3), (6, 5, 7), (21, 8, 12), (40, 33, 20)])
def test_mlw_m_s_3d_call(oracle, dev, shape):
    s = smooth_state_3d(*shape, seed=sum(shape) + 2)
    dx, dxm, dy, dym, dz, dzm = grid3(*shape)
    x, xm = radii(shape[0], 0.4, dx)
    y, ym = polar(shape[1], 0.6, dy)
    a = {k: v.copy(order="F") for k, v in s.items()}
    oracle.mlw_m_s_3d(*[a[k] for k in F3], 1e-3, 3.0, 5 / 3, x, xm, dx, dxm, y, ym, dy, dym, dz, dzm)
    d = {k: to_dev(v, dev) for k, v in s.items()}
    cans.mlw_m_s_3d(*[d[k] for k in F3], 1e-3, 3.0, 5 / 3,
                    *[T(v, dev) for v in (x, xm, dx, dxm, y, ym, dy, dym, dz, dzm)])
    torch.cuda.synchronize()
    for k in F3:
        assert_bitwise(to_host(d[k]), a[k], k)


def test_mlw_m_s_3d_pole_face_call(oracle, dev):
    shape = (8, 9, 6)
    s = smooth_state_3d(*shape, seed=29)
    dx, dxm, dy, dym, dz, dzm = grid3(*shape)
    x, xm = radii(shape[0], 0.5, dx)
    y, ym = polar(shape[1], 0.0, dy)
    ym[0] = 0.0
    a = {k: v.copy(order="F") for k, v in s.items()}
    oracle.mlw_m_s_3d(*[a[k] for k in F3], 1e-3, 3.0, 5 / 3, x, xm, dx, dxm, y, ym, dy, dym, dz, dzm)
    d = {k: to_dev(v, dev) for k, v in s.items()}
    cans.mlw_m_s_3d(*[d[k] for k in F3], 1e-3, 3.0, 5 / 3,
                    *[T(v, dev) for v in (x, xm, dx, dxm, y, ym, dy, dym, dz, dzm)])
    torch.cuda.synchronize()
    for k in F3:
        assert_bitwise(to_host(d[k]), a[k], k)
    assert not np.isfinite(a["ro"][1:-1, 1, 1:-1]).all()


def test_cfl_m_s_3d_call(oracle, dev):
    shape = (30, 21, 17)
    s = smooth_state_3d(*shape, seed=43)
    dx, _, dy, _, dz, _ = grid3(*shape)
    x, _ = radii(shape[0], 0.3, dx)
    y, _ = polar(shape[1], 0.4, dy)
    ref = oracle.cfl_m_s_3d(0.4, 1e-10, 5 / 3, *[s[k] for k in F3], x, dx, y, dy, dz)
    d = [to_dev(s[k], dev) for k in F3]
    got = cans.cfl_m_s_3d(0.4, 1e-10, 5 / 3, *d, *[T(v, dev) for v in (x, dx, y, dy, dz)])
    assert got[0] == ref[0] and got[1] == ref[1]


def test_mlw_m_s_3d_resident_loop(oracle, dev):
    """10 steps of {cfl_m_s; t=t+dt; mlw_m_s; bdper x, y, z}: the spherical
    engine as a resident run (faces periodic, a test of the loop only)."""
    shape, margin, nsteps = (22, 18, 14), 1, 10
    s = smooth_state_3d(*shape, seed=47)
    dx, dxm, dy, dym, dz, dzm = grid3(*shape)
    x, xm = radii(shape[0], 0.6, dx)
    y, ym = polar(shape[1], 0.5, dy)
    for dr in (0, 1, 2):
        for k in F3:
            oracle.bnd_3d(dr, 7, margin, margin, s[k])
    a = {k: v.copy(order="F") for k, v in s.items()}
    t = 0.0
    for _ in range(nsteps):
        dt, merr, _ = oracle.cfl_m_s_3d(0.4, 1e-10, 5 / 3, *[a[k] for k in F3], x, dx, y, dy, dz)
        t = t + dt
        oracle.mlw_m_s_3d(*[a[k] for k in F3], dt, 3.0, 5 / 3, x, xm, dx, dxm, y, ym, dy, dym, dz, dzm)
        for dr in (0, 1, 2):
            for k in F3:
                oracle.bnd_3d(dr, 7, margin, margin, a[k])
    st = [to_dev(s[k], dev) for k in F3]
    grid = {k: T(v, dev) for k, v in zip(("dx", "dxm", "dy", "dym", "dz", "dzm"), (dx, dxm, dy, dym, dz, dzm))}
    geo = T(np.concatenate([x, xm, y, ym]), dev)
    run = cans.Run(3, cans.ENGINE_MLWS, st, grid, 5 / 3, 0.4, qav=3.0, margin=margin, bc=7, bx1d=geo)
    run.reset()
    run.steps(nsteps)
    cur, info = run.state()
    torch.cuda.synchronize()
    assert info["nexec"] == nsteps and info["t"] == t
    for k, g in zip(F3, cur):
        assert_bitwise(to_host(g), a[k], k)


# ---------------------------------------------------------------- mlw_m_s, cfl_m_s, 2D (r, theta)
from tests.test_oracle_m3 import F2S, state_s2  # noqa: E402


@pytest.mark.parametrize("ix,jx", [(5, 6), (33, 9), (70, 51), (130, 71)])
def test_mlw_m_s_2d_call(oracle, dev, ix, jx):
    s, x, xm, dx, dxm, y, ym, dy, dym = state_s2(ix, jx, ix + jx)
    a = {k: v.copy(order="F") for k, v in s.items()}
    oracle.mlw_m_s_2d(*[a[k] for k in F2S], 1e-3, 3.0, 5 / 3, x, xm, dx, dxm, y, ym, dy, dym)
    d = {k: to_dev(v, dev) for k, v in s.items()}
    cans.mlw_m_s_2d(*[d[k] for k in F2S], 1e-3, 3.0, 5 / 3, *[T(v, dev) for v in (x, xm, dx, dxm, y, ym, dy, dym)])
    torch.cuda.synchronize()
    for k in F2S:
        assert_bitwise(to_host(d[k]), a[k], k)


def test_mlw_m_s_2d_pole_face_call(oracle, dev):
    ix, jx = 20, 16
    s, x, xm, dx, dxm, y, ym, dy, dym = state_s2(ix, jx, 5, t0=0.0)
    ym[0] = 0.0
    a = {k: v.copy(order="F") for k, v in s.items()}
    oracle.mlw_m_s_2d(*[a[k] for k in F2S], 1e-3, 3.0, 5 / 3, x, xm, dx, dxm, y, ym, dy, dym)
    d = {k: to_dev(v, dev) for k, v in s.items()}
    cans.mlw_m_s_2d(*[d[k] for k in F2S], 1e-3, 3.0, 5 / 3, *[T(v, dev) for v in (x, xm, dx, dxm, y, ym, dy, dym)])
    torch.cuda.synchronize()
    for k in F2S:
        assert_bitwise(to_host(d[k]), a[k], k)


def test_cfl_m_s_2d_call(oracle, dev):
    ix, jx = 61, 47
    s, x, xm, dx, dxm, y, ym, dy, dym = state_s2(ix, jx, 12)
    f6 = ("ro", "pr", "vx", "vy", "bx", "by")
    ref = oracle.cfl_m_s_2d(0.4, 1e-10, 5 / 3, *[s[k] for k in f6], x, dx, dy)
    got = cans.cfl_m_s_2d(0.4, 1e-10, 5 / 3, *[to_dev(s[k], dev) for k in f6], T(x, dev), T(dx, dev), T(dy, dev))
    assert got[0] == ref[0] and got[1] == ref[1]


def test_mlw_m_s_2d_resident_loop(oracle, dev):
    """20 steps of {cfl_m_s; t=t+dt; mlw_m_s; bdper x, y} as a resident run."""
    ix, jx, margin, nsteps = 50, 40, 1, 20
    s, x, xm, dx, dxm, y, ym, dy, dym = state_s2(ix, jx, 21)
    for dr in (0, 1):
        for k in F2S:
            oracle.bnd_2d(dr, 7, margin, margin, s[k])
    a = {k: v.copy(order="F") for k, v in s.items()}
    t = 0.0
    for _ in range(nsteps):
        dt, merr, _ = oracle.cfl_m_s_2d(0.4, 1e-10, 5 / 3, *[a[k] for k in ("ro", "pr", "vx", "vy", "bx", "by")],
                                        x, dx, dy)
        t = t + dt
        oracle.mlw_m_s_2d(*[a[k] for k in F2S], dt, 3.0, 5 / 3, x, xm, dx, dxm, y, ym, dy, dym)
        for dr in (0, 1):
            for k in F2S:
                oracle.bnd_2d(dr, 7, margin, margin, a[k])
    st = [to_dev(s[k], dev) for k in F2S]
    grid = {"dx": T(dx, dev), "dxm": T(dxm, dev), "dy": T(dy, dev), "dym": T(dym, dev)}
    geo = T(np.concatenate([x, xm, y, ym]), dev)
    run = cans.Run(2, cans.ENGINE_MLWS, st, grid, 5 / 3, 0.4, qav=3.0, margin=margin, bc=7, bx1d=geo)
    run.reset()
    run.steps(nsteps)
    cur, info = run.state()
    torch.cuda.synchronize()
    assert info["nexec"] == nsteps and info["t"] == t
    for k, g in zip(F2S, cur):
        assert_bitwise(to_host(g), a[k], k)
