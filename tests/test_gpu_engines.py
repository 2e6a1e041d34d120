"""GPU parity of the Roe-MUSCL (1D/2D/3D), 3D MLW and HLLD engines, the 3D
cfl_m/bnd and the resident runs of every engine, against the C restatement
of the reference.  FP64 results must be bit-identical.

The Roe and HLLD engines read half-step values of the outermost cell that the
reference never sets, so their first and last interior cells are only
defined for margin >= 3 (the reference drivers use 3); one-call comparisons
therefore cover cells [4, n-3] (the oracle marks the others NaN), and the
resident runs use margin 3, after which the boundary fill defines every cell.
"""
import numpy as np
import pytest

from tests.helpers import (assert_bitwise, grid3, rel_l1, smooth_state_1d, smooth_state_2d, smooth_state_3d,
                           stretched_grid, to_dev, to_host)
from utrip_amd import problems as P

pytestmark = pytest.mark.gpu

F1 = ("ro", "pr", "vx", "vy", "by")
F2 = ("ro", "pr", "vx", "vy", "bx", "by", "az")
F3 = ("ro", "pr", "vx", "vy", "vz", "bx", "by", "bz")


def case(*a):
    print("CASE", *a, flush=True)


def inner(a, lo, dims):
    sl = tuple(slice(lo, n - lo) if n > 1 else slice(None) for n in a.shape[:dims]) + (slice(None),) * (a.ndim - dims)
    return a[sl]


# ------------------------------------------------------------------ Roe one call
@pytest.mark.parametrize("ix", [9, 80, 300])
def test_roe1d_one_call(dev, oracle, ix):
    from utrip_amd import cans
    case("roe1d", ix)
    s = smooth_state_1d(ix, seed=ix)
    _, dx, _ = stretched_grid(ix, seed=ix)
    ref = {k: v.copy() for k, v in s.items()}
    oracle.roe_m_1d(ref["ro"], ref["pr"], ref["vx"], ref["vy"], ref["by"], ref["bx"], ref["bxm"], 1.1e-3, 5 / 3, dx)
    t = {k: to_dev(v, dev) for k, v in s.items()}
    cans.roe_m_1d(t["ro"], t["pr"], t["vx"], t["vy"], t["by"], t["bx"], t["bxm"], 1.1e-3, 5 / 3, to_dev(dx, dev))
    for k in F1:
        assert_bitwise(to_host(t[k])[3:-3], ref[k][3:-3], k)


@pytest.mark.parametrize("ix,jx", [(9, 9), (40, 33), (70, 45)])
def test_roe2d_one_call(dev, oracle, ix, jx):
    from utrip_amd import cans
    case("roe2d", ix, jx)
    s = smooth_state_2d(ix, jx, seed=ix + jx)
    _, dx, _ = stretched_grid(ix, 2)
    _, dy, _ = stretched_grid(jx, 3)
    ref = {k: v.copy(order="F") for k, v in s.items()}
    oracle.roe_m_2d(*[ref[k] for k in F2], 9e-4, 5 / 3, dx, dy)
    t = {k: to_dev(v, dev) for k, v in s.items()}
    cans.roe_m_2d(*[t[k] for k in F2], 9e-4, 5 / 3, to_dev(dx, dev), to_dev(dy, dev))
    for k in F2:
        assert_bitwise(inner(to_host(t[k]), 3, 2), inner(ref[k], 3, 2), k)


@pytest.mark.parametrize("shape", [(9, 9, 9), (14, 12, 11), (33, 10, 12)])
def test_roe3d_one_call(dev, oracle, shape):
    from utrip_amd import cans
    case("roe3d", shape)
    s = smooth_state_3d(*shape, seed=sum(shape))
    dx, _, dy, _, dz, _ = grid3(*shape)
    ref = {k: v.copy(order="F") for k, v in s.items()}
    oracle.roe_m_3d(*[ref[k] for k in F3], 8e-4, 5 / 3, dx, dy, dz)
    t = {k: to_dev(v, dev) for k, v in s.items()}
    cans.roe_m_3d(*[t[k] for k in F3], 8e-4, 5 / 3, *[to_dev(v, dev) for v in (dx, dy, dz)])
    for k in F3:
        assert_bitwise(inner(to_host(t[k]), 3, 3), inner(ref[k], 3, 3), k)


# ------------------------------------------------------------------ 3D MLW one call
# Path 1 (default) is the z-marching kernel on 16x16 column tiles (14x14
# updated cells): (17, 9, 70) and (40, 7, 9) cut the last x tile short,
# (31, 16, 34) has three x tiles and one y tile of exactly 14, (47, 33, 12)
# three by three tiles; path 0 is the three kernels over HBM scratch.
@pytest.mark.parametrize("path", [1, 0])
@pytest.mark.parametrize("shape", [(3, 3, 3), (13, 11, 10), (40, 7, 9), (17, 9, 70), (31, 16, 34), (16, 8, 66),
                                   (47, 33, 12)])
def test_mlw3d_one_call(dev, oracle, shape, path):
    from utrip_amd import cans
    case("mlw3d", shape, "path", path)
    s = smooth_state_3d(*shape, seed=7 + sum(shape))
    dx, dxm, dy, dym, dz, dzm = grid3(*shape)
    ref = {k: v.copy(order="F") for k, v in s.items()}
    oracle.mlw_m_3d(*[ref[k] for k in F3], 1.2e-3, 3.0, 5 / 3, dx, dxm, dy, dym, dz, dzm)
    t = {k: to_dev(v, dev) for k, v in s.items()}
    cans.set_mlw3d_path(path)
    try:
        cans.mlw_m_3d(*[t[k] for k in F3], 1.2e-3, 3.0, 5 / 3, *[to_dev(v, dev) for v in (dx, dxm, dy, dym, dz, dzm)])
    finally:
        cans.set_mlw3d_path(1)
    for k in F3:
        assert_bitwise(to_host(t[k]), ref[k], k)


# ------------------------------------------------------------------ HLLD one call
@pytest.mark.parametrize("shape,mf", [((14, 12, 11), (1, 1, 1)), ((60, 1, 1), (1, 0, 0)), ((20, 17, 1), (1, 1, 0)),
                                      ((9, 9, 9), (1, 1, 1))])
def test_hlld_one_call(dev, oracle, shape, mf):
    from utrip_amd import cans
    case("hlld", shape, mf)
    s = smooth_state_3d(*shape, seed=11 + sum(shape))
    dx, _, dy, _, dz, _ = grid3(*shape)
    ref = {k: v.copy(order="F") for k, v in s.items()}
    oracle.hlld_engine(*[ref[k] for k in F3], 5 / 3, 1.0, 7e-4, dx, dy, dz, mf)
    t = {k: to_dev(v, dev) for k, v in s.items()}
    cans.hlld_engine_3d(*[t[k] for k in F3], 5 / 3, 1.0, 7e-4, *[to_dev(v, dev) for v in (dx, dy, dz)], mfdim=mf)
    for k in F3:
        assert_bitwise(inner(to_host(t[k]), 3, 3), inner(ref[k], 3, 3), k)


# ------------------------------------------------------------------ 3D cfl and bnd
def test_cfl3d(dev, oracle):
    from utrip_amd import cans
    s = smooth_state_3d(15, 12, 9, seed=3)
    dx, _, dy, _, dz, _ = grid3(15, 12, 9)
    s["pr"][4, 5, 6] = np.nan
    dto, merro, _ = oracle.cfl_m_3d(0.4, 1e-10, 5 / 3, *[s[k] for k in F3], dx, dy, dz)
    t = {k: to_dev(v, dev) for k, v in s.items()}
    dtg, merrg = cans.cfl_m_3d(0.4, 1e-10, 5 / 3, *[t[k] for k in F3], *[to_dev(v, dev) for v in (dx, dy, dz)])
    assert dtg == dto and merrg == merro == 0
    dtg, merrg = cans.cfl_m_3d(0.4, 1e3, 5 / 3, *[t[k] for k in F3], *[to_dev(v, dev) for v in (dx, dy, dz)])
    assert merrg == 9001


@pytest.mark.parametrize("kind", range(8))
def test_bnd3d_kinds(dev, oracle, kind):
    from utrip_amd import cans
    rng = np.random.default_rng(kind)
    shape = (9, 8, 7)
    for d in range(3):
        _, _, dm = stretched_grid(shape[d], seed=d)
        for margin in (1, 2):
            sides = [(margin, margin)] if kind == 7 else [(0, margin), (1, margin)]
            for mbnd, mg in sides:
                q = np.asfortranarray(rng.standard_normal(shape))
                r = q.copy(order="F")
                oracle.bnd_3d(d, kind, mbnd, mg, r, 0.37, dm)
                t = to_dev(q, dev)
                cans.bnd_3d(d, kind, mbnd, mg, t, 0.37, to_dev(dm, dev))
                assert_bitwise(to_host(t), r, f"3d dir {d} kind {kind} side {mbnd} margin {margin}")


# ------------------------------------------------------------------ resident runs
def oracle_loop(oracle, dim, engine, fields, grid, pp, nsteps, bc_kind, margin, bx=None, bxm=None, mu=1.0):
    """The reference main loop on the CPU: cfl_m, engine, bnd (all fields,
    x then y then z faces, one boundary kind everywhere)."""
    f = {k: v.copy(order="F") for k, v in fields.items()}
    names = {1: F1, 2: F2, 3: F3}[dim]
    dts = []
    for _ in range(nsteps):
        if dim == 1:
            dt, merr, _ = oracle.cfl_m_1d(pp["safety"], pp["dtmin"], pp["gm"], bx, f["ro"], f["pr"], f["vx"], f["vy"],
                                          f["by"], grid["dx"])
        elif dim == 2:
            dt, merr, _ = oracle.cfl_m_2d(pp["safety"], pp["dtmin"], pp["gm"], *[f[k] for k in F2[:6]], grid["dx"],
                                          grid["dy"])
        else:
            dt, merr, _ = oracle.cfl_m_3d(pp["safety"], pp["dtmin"], pp["gm"], *[f[k] for k in F3], grid["dx"],
                                          grid["dy"], grid["dz"])
        assert merr == 0
        if engine == "mlw":
            if dim == 1:
                oracle.mlw_m_1d(*[f[k] for k in F1], bx, bxm, dt, pp["qav"], pp["gm"], grid["dx"], grid["dxm"])
            elif dim == 2:
                oracle.mlw_m_2d(*[f[k] for k in F2], dt, pp["qav"], pp["gm"], grid["dx"], grid["dxm"], grid["dy"],
                                grid["dym"])
            else:
                oracle.mlw_m_3d(*[f[k] for k in F3], dt, pp["qav"], pp["gm"], grid["dx"], grid["dxm"], grid["dy"],
                                grid["dym"], grid["dz"], grid["dzm"])
        elif engine == "roe":
            if dim == 1:
                oracle.roe_m_1d(*[f[k] for k in F1], bx, bxm, dt, pp["gm"], grid["dx"])
            elif dim == 2:
                oracle.roe_m_2d(*[f[k] for k in F2], dt, pp["gm"], grid["dx"], grid["dy"])
            else:
                oracle.roe_m_3d(*[f[k] for k in F3], dt, pp["gm"], grid["dx"], grid["dy"], grid["dz"])
        else:
            oracle.hlld_engine(*[f[k] for k in F3], pp["gm"], mu, dt, grid["dx"], grid["dy"], grid["dz"], (1, 1, 1))
        for d in range(dim):
            dm = grid[("dxm", "dym", "dzm")[d]]
            for k in names:
                if bc_kind == 7:
                    (oracle.bnd_1d(7, margin, margin, f[k]) if dim == 1 else
                     oracle.bnd_2d(d, 7, margin, margin, f[k]) if dim == 2 else
                     oracle.bnd_3d(d, 7, margin, margin, f[k]))
                else:
                    for side in (0, 1):
                        (oracle.bnd_1d(bc_kind, side, margin, f[k], 0.0, dm) if dim == 1 else
                         oracle.bnd_2d(d, bc_kind, side, margin, f[k], 0.0, dm) if dim == 2 else
                         oracle.bnd_3d(d, bc_kind, side, margin, f[k], 0.0, dm))
        dts.append(dt)
    return f, dts


def run_gpu(dev, dim, engine, fields, grid, pp, nsteps, bc_kind, margin, bx=None, bxm=None, mixed=False, mu=1.0):
    from utrip_amd import cans
    names = {1: F1, 2: F2, 3: F3}[dim]
    st = [to_dev(fields[k], dev) for k in names]
    gd = {k: to_dev(v, dev) for k, v in grid.items() if k in ("dx", "dxm", "dy", "dym", "dz", "dzm")}
    eng = {"mlw": cans.ENGINE_MLW, "roe": cans.ENGINE_ROE, "hlld": cans.ENGINE_HLLD}[engine]
    run = cans.Run(dim, eng, st, gd, pp["gm"], pp["safety"], qav=pp.get("qav", 3.0), mu=mu, margin=margin,
                   bc=bc_kind, mixed=mixed, bx1d=to_dev(bx, dev) if bx is not None else None,
                   bxm1d=to_dev(bxm, dev) if bxm is not None else None)
    run.reset(dtmin=pp["dtmin"])
    run.steps(nsteps // 2)
    run.steps(nsteps - nsteps // 2)
    cur, info = run.state()
    return {k: to_host(t) for k, t in zip(names, cur)}, info


@pytest.mark.parametrize("engine", ["mlw", "roe"])
def test_run_1d_shock_tube(dev, oracle, engine):
    case("run 1d", engine)
    margin = 3
    pb = P.shock_tube_1d(nx=256, margin=margin)
    f = {k: pb.fields[k] for k in F1}
    ref, dts = oracle_loop(oracle, 1, engine, f, pb.grid, pb.params, 60, 0, margin, pb.fields["bx"], pb.fields["bxm"])
    got, info = run_gpu(dev, 1, engine, f, pb.grid, pb.params, 60, 0, margin, pb.fields["bx"], pb.fields["bxm"])
    assert info["nexec"] == 60 and info["dt"] == dts[-1]
    for k in F1:
        assert_bitwise(got[k], ref[k], k)


@pytest.mark.parametrize("engine,margin", [("mlw", 1), ("roe", 3)])
def test_run_2d_ot(dev, oracle, engine, margin):
    case("run 2d OT", engine)
    pb = P.orszag_tang_2d(n=48, margin=margin)
    ref, dts = oracle_loop(oracle, 2, engine, pb.fields, pb.grid, pb.params, 12, 7, margin)
    got, info = run_gpu(dev, 2, engine, pb.fields, pb.grid, pb.params, 12, 7, margin)
    assert info["nexec"] == 12 and info["dt"] == dts[-1]
    for k in F2:
        assert_bitwise(got[k], ref[k], k)


@pytest.mark.parametrize("shape,margin", [((13, 17, 11), 1), ((20, 9, 12), 2), ((10, 14, 19), 3), ((21, 13), 2),
                                          ((23, 9), 3), ((40,), 2)])
def test_run_periodic_fill_shapes(dev, oracle, shape, margin):
    """Periodic MLW runs whose ghost fill is the one-pass periodic fill (every
    field periodic on every face) on uneven 1D/2D/3D boxes and margins 1-3,
    against the reference's bdperx, bdpery, bdperz one after the other (the
    oracle loop), bit for bit with the dt sequence; the initial ghosts are not
    images, so the first step's cfl pass reads them as the reference does."""
    case("run periodic fill", shape, "margin", margin)
    dim = len(shape)
    if dim == 3:
        f = smooth_state_3d(*shape, seed=sum(shape))
        dx, dxm, dy, dym, dz, dzm = grid3(*shape)
        grid = dict(dx=dx, dxm=dxm, dy=dy, dym=dym, dz=dz, dzm=dzm)
        fields, bx, bxm = {k: f[k] for k in F3}, None, None
    elif dim == 2:
        f = smooth_state_2d(*shape, seed=sum(shape))
        _, dx, dxm = stretched_grid(shape[0], 5)
        _, dy, dym = stretched_grid(shape[1], 6)
        grid = dict(dx=dx, dxm=dxm, dy=dy, dym=dym)
        fields, bx, bxm = {k: f[k] for k in F2}, None, None
    else:
        f = smooth_state_1d(shape[0], seed=3)
        _, dx, dxm = stretched_grid(shape[0], 7)
        grid = dict(dx=dx, dxm=dxm)
        fields, bx, bxm = {k: f[k] for k in F1}, f["bx"], f["bxm"]
    pp = dict(gm=5.0 / 3.0, safety=0.4, qav=3.0, dtmin=1e-10)
    ref, dts = oracle_loop(oracle, dim, "mlw", fields, grid, pp, 6, 7, margin, bx, bxm)
    got, info = run_gpu(dev, dim, "mlw", fields, grid, pp, 6, 7, margin, bx, bxm)
    assert info["nexec"] == 6 and info["merr"] == 0 and info["dt"] == dts[-1]
    for k in {1: F1, 2: F2, 3: F3}[dim]:
        assert_bitwise(got[k], ref[k], k)


@pytest.mark.parametrize("n,nsteps", [(40, 12), (724, 4)])
@pytest.mark.parametrize("axis", ["x", "y"])
def test_run_2d_roe_periodic_ghost_cfl(dev, oracle, axis, n, nsteps):
    """Periodic 2D Roe run (after the first step of a call the stage-1 x faces
    take the step's cfl_m and the next kernel advances the clock) on
    stretched grids whose smallest spacing on one axis lies in the two ghost
    cells at its ends: the reference's cfl_m takes dt from a ghost image.
    724^2 takes the fused y marches, 40^2 the separate passes; two calls."""
    from utrip_amd.problems import Problem
    case("run 2d roe periodic ghost cfl", n, "axis", axis)
    margin = 3
    base = P.orszag_tang_2d(n=n, margin=margin)
    ix, jx = base.fields["ro"].shape
    grid = dict(base.grid)
    for nm, m, seed in (("x", ix, 31), ("y", jx, 32)):
        _, d, dm = stretched_grid(m, seed=seed)
        if nm == axis:
            d[0], d[m - 1] = 0.3 * d.min(), 0.35 * d.min()
        grid["d" + nm], grid["d" + nm + "m"] = d, dm
    pb = Problem(base.name, base.fields, grid, base.params)
    ref, dts = oracle_loop(oracle, 2, "roe", pb.fields, grid, pb.params, nsteps, 7, margin)
    got, info = run_gpu(dev, 2, "roe", pb.fields, grid, pb.params, nsteps, 7, margin)
    assert info["nexec"] == nsteps and info["merr"] == 0 and info["dt"] == dts[-1]
    for k in F2:
        assert_bitwise(got[k], ref[k], k)


@pytest.mark.parametrize("engine,margin", [("mlw", 2), ("roe", 3), ("hlld", 3)])
def test_run_3d_blast(dev, oracle, engine, margin):
    case("run 3d blast", engine)
    pb = P.blast_3d(n=14, margin=margin, gaussian_b=(engine != "hlld"), r0=0.25)
    ref, dts = oracle_loop(oracle, 3, engine, pb.fields, pb.grid, pb.params, 6, 7, margin)
    got, info = run_gpu(dev, 3, engine, pb.fields, pb.grid, pb.params, 6, 7, margin)
    assert info["nexec"] == 6 and info["dt"] == dts[-1]
    for k in F3:
        assert_bitwise(got[k], ref[k], k)


@pytest.mark.parametrize("axis", ["x", "y", "z"])
@pytest.mark.parametrize("engine,margin", [("mlw", 2), ("mlw", 3), ("roe", 3), ("hlld", 3)])
def test_run_3d_periodic_ghost_cfl(dev, oracle, engine, margin, axis):
    """Periodic 3D runs whose engine advances the clock and takes the next
    step's cfl_m (the MLW marching kernel; the Roe step with its fused z
    marches, forced here by a low wave target; HLLD) on stretched grids whose
    smallest spacing on one axis lies in the two ghost cells at its ends: the
    reference's cfl_m takes dt from a ghost image, which the kernels fold into
    its interior cell's term.  The run is issued in two calls (the second
    call's first step takes cfl_m in a pass of its own)."""
    from utrip_amd import cans
    from utrip_amd.problems import Problem
    case("run 3d periodic ghost cfl", engine, "margin", margin, "axis", axis)
    base = P.blast_3d(n=14, margin=margin, gaussian_b=(engine != "hlld"), r0=0.25)
    ix, jx, kx = base.fields["ro"].shape
    grid = dict(base.grid)
    for nm, n, seed in (("x", ix, 21), ("y", jx, 22), ("z", kx, 23)):
        _, d, dm = stretched_grid(n, seed=seed)
        if nm == axis:
            d[0], d[n - 1] = 0.3 * d.min(), 0.35 * d.min()
        grid["d" + nm], grid["d" + nm + "m"] = d, dm
    pb = Problem(base.name, base.fields, grid, base.params)
    ref, dts = oracle_loop(oracle, 3, engine, pb.fields, grid, pb.params, 8, 7, margin)
    cans.set_march_waves(8 if engine == "roe" else 32768)
    try:
        got, info = run_gpu(dev, 3, engine, pb.fields, grid, pb.params, 8, 7, margin)
    finally:
        cans.set_march_waves(32768)
    assert info["nexec"] == 8 and info["merr"] == 0 and info["dt"] == dts[-1]
    for k in F3:
        assert_bitwise(got[k], ref[k], k)
    if engine == "roe":
        # the Roe step takes the fused cfl_m with FP32 fluxes: the fused z
        # marches (low wave target) against the separate passes and cfl pass
        cans.set_march_waves(8)
        try:
            got32, info32 = run_gpu(dev, 3, engine, pb.fields, grid, pb.params, 8, 7, margin, mixed=True)
        finally:
            cans.set_march_waves(32768)
        ref32, info_ref32 = run_gpu(dev, 3, engine, pb.fields, grid, pb.params, 8, 7, margin, mixed=True)
        assert info32["nexec"] == 8 and info32["dt"] == info_ref32["dt"]
        for k in F3:
            assert_bitwise(got32[k], ref32[k], "fp32-flux " + k)


@pytest.mark.parametrize("engine,margin,mixed", [("mlw", 2, False), ("roe", 3, False), ("roe", 3, True),
                                                 ("hlld", 3, False)])
def test_run_3d_periodic_stops(dev, oracle, engine, margin, mixed):
    """The clock carried by the periodic 3D engines' kernels (FP64 Roe: the
    x faces take the step's cfl_m, the y faces advance the clock; FP32-flux
    Roe on its fused z marches, forced by a low wave target): the t < tend
    loop test ends the run after the step that reaches tend (later steps are
    no-ops), and dt < dtmin sets merr = 9001 with the state unchanged."""
    from utrip_amd import cans
    case("run 3d periodic, tend and dtmin stops", engine, "fp32-flux" if mixed else "fp64")
    pb = P.blast_3d(n=14, margin=margin, gaussian_b=(engine != "hlld"), r0=0.25)
    pp = pb.params
    if mixed:
        # the reference: the same FP32-flux run on the separate passes and cfl pass
        _, info3 = run_gpu(dev, 3, engine, pb.fields, pb.grid, pp, 3, 7, margin, mixed=True)
        t = info3["t"]
        ref, _ = run_gpu(dev, 3, engine, pb.fields, pb.grid, pp, 4, 7, margin, mixed=True)
    else:
        ref, dts = oracle_loop(oracle, 3, engine, pb.fields, pb.grid, pp, 4, 7, margin)
        t = 0.0
        for d in dts[:3]:
            t = t + d
    st = [to_dev(pb.fields[k], dev) for k in F3]
    gd = {k: to_dev(v, dev) for k, v in pb.grid.items() if k in ("dx", "dxm", "dy", "dym", "dz", "dzm")}
    eng = {"mlw": cans.ENGINE_MLW, "roe": cans.ENGINE_ROE, "hlld": cans.ENGINE_HLLD}[engine]
    cans.set_march_waves(8 if engine == "roe" else 32768)
    try:
        run = cans.Run(3, eng, st, gd, pp["gm"], pp["safety"], qav=pp.get("qav", 3.0), margin=margin, bc=7,
                       mixed=mixed)
        run.reset(tend=t * (1 + 1e-12), dtmin=pp["dtmin"])   # reached after step 4
        run.steps(3)
        run.steps(5)
        cur, info = run.state()
        assert info["nexec"] == 4 and info["ns"] == 4 and info["merr"] == 0
        for k, tt in zip(F3, cur):
            assert_bitwise(to_host(tt), ref[k], k)
        st = [to_dev(pb.fields[k], dev) for k in F3]
        run = cans.Run(3, eng, st, gd, pp["gm"], pp["safety"], qav=pp.get("qav", 3.0), margin=margin, bc=7,
                       mixed=mixed)
        run.reset(dtmin=1.0)
        run.steps(3)
        cur, info = run.state()
    finally:
        cans.set_march_waves(32768)
    assert info["merr"] == 9001 and info["nexec"] == 0 and info["ns"] == 1
    for k, tt in zip(F3, cur):
        assert_bitwise(to_host(tt), pb.fields[k], k)


# Longer runs of the marching and fused kernels against the oracle: the 3D
# MLW march on several column tiles and z-segments, the 3D Roe and HLLD with
# the stage updates fused into their z sweeps (several segments per column
# at the default wave target), the 2D Roe with the fused y marches (taken on
# grids of at least 512K y faces: 724^2 is just above; 80 segments of 9
# faces per column): every field and the dt sequence, bit for bit.
@pytest.mark.parametrize("dim,engine,n,margin,nsteps", [(3, "mlw", 30, 2, 40), (3, "roe", 26, 3, 30),
                                                        (3, "hlld", 26, 3, 30), (2, "roe", 724, 3, 16)])
def test_longer_runs_of_the_fused_paths(dev, oracle, dim, engine, n, margin, nsteps):
    from utrip_amd import cans
    case("longer run", dim, engine, n, nsteps)
    if dim == 3:
        pb = P.blast_3d(n=n, margin=margin, gaussian_b=(engine != "hlld"), r0=0.25)
        names = F3
    else:
        pb = P.orszag_tang_2d(n=n, margin=margin)
        names = F2
    ref, dts = oracle_loop(oracle, dim, engine, pb.fields, pb.grid, pb.params, nsteps, 7, margin)
    # 3D: a low wave target, so that the march segments are long (the 3D Roe
    # takes its fused z march on segments of >= 8 faces)
    cans.set_march_waves(8 if dim == 3 else 32768)
    try:
        got, info = run_gpu(dev, dim, engine, pb.fields, pb.grid, pb.params, nsteps, 7, margin)
    finally:
        cans.set_march_waves(32768)
    assert info["nexec"] == nsteps and info["dt"] == dts[-1]
    for k in names:
        assert_bitwise(got[k], ref[k], k)


@pytest.mark.parametrize("waves", [1, 8])
@pytest.mark.parametrize("engine", ["roe", "hlld"])
def test_run_3d_blast_long_segments(dev, oracle, engine, waves):
    """The marching face kernels with long segments: at this size the default
    wave target gives one face per thread, so force whole columns (1 wave)
    and a few segments per column (8 waves); bit-exact either way."""
    from utrip_amd import cans
    case("run 3d blast, march waves", engine, waves)
    margin = 3
    pb = P.blast_3d(n=14, margin=margin, gaussian_b=(engine != "hlld"), r0=0.25)
    ref, dts = oracle_loop(oracle, 3, engine, pb.fields, pb.grid, pb.params, 6, 7, margin)
    cans.set_march_waves(waves)
    try:
        got, info = run_gpu(dev, 3, engine, pb.fields, pb.grid, pb.params, 6, 7, margin)
        got32, _ = run_gpu(dev, 3, engine, pb.fields, pb.grid, pb.params, 6, 7, margin, mixed=True)
    finally:
        cans.set_march_waves(32768)
    assert info["nexec"] == 6 and info["dt"] == dts[-1]
    for k in F3:
        assert_bitwise(got[k], ref[k], k)
    ref32, _ = run_gpu(dev, 3, engine, pb.fields, pb.grid, pb.params, 6, 7, margin, mixed=True)
    for k in F3:
        assert_bitwise(got32[k], ref32[k], "fp32-flux " + k)


def test_march_waves_rejects_zero(dev):
    from utrip_amd import cans
    with pytest.raises(cans.CansError):
        cans.set_march_waves(0)


@pytest.mark.parametrize("engine", ["roe", "hlld"])
def test_mixed_precision_3d(dev, oracle, engine):
    """FP32 face fluxes + FP64 state: density within 1e-5 relative L1 of the
    FP64 oracle after a few steps of the blast wave."""
    case("mixed 3d", engine)
    margin = 3
    pb = P.blast_3d(n=16, margin=margin, gaussian_b=(engine != "hlld"), r0=0.25)
    ref, _ = oracle_loop(oracle, 3, engine, pb.fields, pb.grid, pb.params, 5, 7, margin)
    got, info = run_gpu(dev, 3, engine, pb.fields, pb.grid, pb.params, 5, 7, margin, mixed=True)
    err = {k: rel_l1(got[k], ref[k]) for k in F3}
    print("rel L1 mixed vs fp64:", err)
    assert err["ro"] < 1e-5, err
    assert max(err[k] for k in ("ro", "pr", "bx", "by")) < 1e-4, err
