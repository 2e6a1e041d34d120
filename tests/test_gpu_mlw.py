"""GPU parity of the modified Lax-Wendroff engines, cfl_m and bnd against the
C restatement of the reference (bit-identical FP64 is the bar: the kernels
keep the reference's operation order and use no FMA contraction)."""
import numpy as np
import pytest

from tests.helpers import assert_bitwise, smooth_state_1d, smooth_state_2d, stretched_grid, to_dev, to_host
from utrip_amd import problems as P

pytestmark = pytest.mark.gpu

F1 = ("ro", "pr", "vx", "vy", "by")
F2 = ("ro", "pr", "vx", "vy", "bx", "by", "az")


def _print_case(*a):
    print("CASE", *a, flush=True)


# ----------------------------------------------------------------- 1D
@pytest.mark.parametrize("ix", [3, 5, 257, 600])
def test_mlw1d_one_call(dev, oracle, ix):
    from utrip_amd import cans
    _print_case("mlw1d one call", ix)
    s = smooth_state_1d(ix, seed=ix)
    x, dx, dxm = stretched_grid(ix, seed=ix)
    ref = {k: v.copy() for k, v in s.items()}
    oracle.mlw_m_1d(ref["ro"], ref["pr"], ref["vx"], ref["vy"], ref["by"], ref["bx"], ref["bxm"],
                    1.3e-3, 3.0, 5.0 / 3.0, dx, dxm)
    t = {k: to_dev(v, dev) for k, v in s.items()}
    cans.mlw_m_1d(t["ro"], t["pr"], t["vx"], t["vy"], t["by"], t["bx"], t["bxm"], 1.3e-3, 3.0, 5.0 / 3.0,
                  to_dev(dx, dev), to_dev(dxm, dev))
    for k in F1:
        assert_bitwise(to_host(t[k]), ref[k], k)


def _loop_1d_oracle(oracle, pb, nsteps):
    f = {k: v.copy() for k, v in pb.fields.items()}
    g, pp = pb.grid, pb.params
    m = pp["margin"]
    dts = []
    for _ in range(nsteps):
        dt, merr, _i = oracle.cfl_m_1d(pp["safety"], pp["dtmin"], pp["gm"], f["bx"], f["ro"], f["pr"], f["vx"],
                                       f["vy"], f["by"], g["dx"])
        assert merr == 0
        oracle.mlw_m_1d(f["ro"], f["pr"], f["vx"], f["vy"], f["by"], f["bx"], f["bxm"], dt, pp["qav"], pp["gm"],
                        g["dx"], g["dxm"])
        for k in F1:
            oracle.bnd_1d(0, 0, m, f[k])
            oracle.bnd_1d(0, 1, m, f[k])
        dts.append(dt)
    return f, dts


def test_mlw1d_shock_tube_dropin_loop(dev, oracle):
    """Brio-Wu (nx=1024, margin 2, bdfrex): 300 steps of the reference loop
    through the drop-in entry points."""
    from utrip_amd import cans
    _print_case("mlw1d shock tube 300 steps")
    pb = P.shock_tube_1d(nx=1024, margin=2)
    ref, dts = _loop_1d_oracle(oracle, pb, 300)
    g, pp = pb.grid, pb.params
    t = {k: to_dev(v, dev) for k, v in pb.fields.items()}
    dx, dxm = to_dev(g["dx"], dev), to_dev(g["dxm"], dev)
    gdts = []
    for _ in range(300):
        dt, merr = cans.cfl_m_1d(pp["safety"], pp["dtmin"], pp["gm"], t["bx"], t["ro"], t["pr"], t["vx"], t["vy"],
                                 t["by"], dx)
        assert merr == 0
        cans.mlw_m_1d(t["ro"], t["pr"], t["vx"], t["vy"], t["by"], t["bx"], t["bxm"], dt, pp["qav"], pp["gm"],
                      dx, dxm)
        for k in F1:
            cans.bnd_1d("fre", 0, 2, t[k])
            cans.bnd_1d("fre", 1, 2, t[k])
        gdts.append(dt)
    assert gdts == dts
    for k in F1:
        assert_bitwise(to_host(t[k]), ref[k], k)


# ----------------------------------------------------------------- 2D
@pytest.mark.parametrize("ix,jx", [(3, 3), (34, 10), (35, 11), (70, 45), (130, 67)])
def test_mlw2d_one_call(dev, oracle, ix, jx):
    from utrip_amd import cans
    _print_case("mlw2d one call", ix, jx)
    s = smooth_state_2d(ix, jx, seed=ix * 100 + jx)
    _, dx, dxm = stretched_grid(ix, seed=ix)
    _, dy, dym = stretched_grid(jx, seed=jx + 7)
    ref = {k: v.copy(order="F") for k, v in s.items()}
    oracle.mlw_m_2d(*[ref[k] for k in F2], 2.1e-3, 3.0, 5.0 / 3.0, dx, dxm, dy, dym)
    t = {k: to_dev(v, dev) for k, v in s.items()}
    cans.mlw_m_2d(*[t[k] for k in F2], 2.1e-3, 3.0, 5.0 / 3.0, to_dev(dx, dev), to_dev(dxm, dev),
                  to_dev(dy, dev), to_dev(dym, dev))
    for k in F2:
        assert_bitwise(to_host(t[k]), ref[k], k)


def _loop_2d_oracle(oracle, pb, nsteps):
    f = {k: v.copy(order="F") for k, v in pb.fields.items()}
    g, pp = pb.grid, pb.params
    m = pp["margin"]
    dts = []
    for _ in range(nsteps):
        dt, merr, _loc = oracle.cfl_m_2d(pp["safety"], pp["dtmin"], pp["gm"], f["ro"], f["pr"], f["vx"], f["vy"],
                                         f["bx"], f["by"], g["dx"], g["dy"])
        assert merr == 0
        oracle.mlw_m_2d(*[f[k] for k in F2], dt, pp["qav"], pp["gm"], g["dx"], g["dxm"], g["dy"], g["dym"])
        for k in F2:
            oracle.bnd_2d(0, 7, m, m, f[k])
        for k in F2:
            oracle.bnd_2d(1, 7, m, m, f[k])
        dts.append(dt)
    return f, dts


def test_mlw2d_ot_dropin_loop(dev, oracle):
    from utrip_amd import cans
    _print_case("mlw2d OT 96^2 drop-in 20 steps")
    pb = P.orszag_tang_2d(n=96, margin=1)
    ref, dts = _loop_2d_oracle(oracle, pb, 20)
    g, pp = pb.grid, pb.params
    t = {k: to_dev(v, dev) for k, v in pb.fields.items()}
    gd = {k: to_dev(v, dev) for k, v in g.items()}
    gdts = []
    for _ in range(20):
        dt, merr = cans.cfl_m_2d(pp["safety"], pp["dtmin"], pp["gm"], *[t[k] for k in F2[:6]], gd["dx"], gd["dy"])
        cans.mlw_m_2d(*[t[k] for k in F2], dt, pp["qav"], pp["gm"], gd["dx"], gd["dxm"], gd["dy"], gd["dym"])
        for k in F2:
            cans.bnd_2d(0, "per", 1, 1, t[k])
        for k in F2:
            cans.bnd_2d(1, "per", 1, 1, t[k])
        gdts.append(dt)
    assert gdts == dts
    for k in F2:
        assert_bitwise(to_host(t[k]), ref[k], k)


@pytest.mark.parametrize("fused", [True, False])
def test_mlw2d_ot_resident(dev, oracle, fused):
    """Resident loop (device clock, ping-pong buffers); fused = periodic
    ghost fill and next-step cfl inside the engine kernel."""
    from utrip_amd import cans
    _print_case("mlw2d OT 80x72 resident fused=", fused)
    pb = P.orszag_tang_2d(n=72, margin=1)
    # make it non-square: 80 x 72 interior
    pb = P.orszag_tang_2d(n=72, margin=1)
    nsteps = 40
    ref, dts = _loop_2d_oracle(oracle, pb, nsteps)
    g, pp = pb.grid, pb.params
    st = [to_dev(pb.fields[k], dev) for k in F2]
    gd = {k: to_dev(v, dev) for k, v in g.items()}
    bc = None if fused else [cans.BD_PER] * 28
    run = cans.MlwRun2D(st, gd["dx"], gd["dxm"], gd["dy"], gd["dym"], pp["qav"], pp["gm"], pp["safety"],
                        margin=1, bc=bc)
    run.reset(dtmin=pp["dtmin"])
    run.steps(17)
    run.steps(nsteps - 17)
    cur, info = run.state()
    assert info["nexec"] == nsteps and info["ns"] == nsteps and info["merr"] == 0
    assert info["dt"] == dts[-1]
    assert info["t"] == sum_sequential(dts)
    for k, tt in zip(F2, cur):
        assert_bitwise(to_host(tt), ref[k], k)


@pytest.mark.parametrize("margin,axis", [(1, "x"), (2, "x"), (1, "y"), (2, "y")])
def test_mlw2d_resident_periodic_ghost_cfl(dev, oracle, margin, axis):
    """Fused periodic run on a stretched grid whose smallest spacings lie in
    ghost cells (dx(1), dx(ix) or dy(1), dy(jx)): the reference's cfl_m takes
    dt from a ghost image, which the fused kernel folds into one term per
    cell."""
    from utrip_amd import cans
    from utrip_amd.problems import Problem
    _print_case("mlw2d resident periodic ghost cfl, margin", margin, "axis", axis)
    base = P.orszag_tang_2d(n=40, margin=margin)
    ix, jx = base.fields["ro"].shape
    _, dx, dxm = stretched_grid(ix, seed=11)
    _, dy, dym = stretched_grid(jx, seed=12)
    if axis == "x":
        dx[0], dx[ix - 1] = 0.3 * dx.min(), 0.35 * dx.min()
    else:
        dy[0], dy[jx - 1] = 0.35 * dy.min(), 0.3 * dy.min()
    grid = dict(base.grid, dx=dx, dxm=dxm, dy=dy, dym=dym)
    pb = Problem(base.name, base.fields, grid, base.params)
    nsteps = 12
    ref, dts = _loop_2d_oracle(oracle, pb, nsteps)
    pp = pb.params
    st = [to_dev(pb.fields[k], dev) for k in F2]
    gd = {k: to_dev(v, dev) for k, v in grid.items()}
    run = cans.MlwRun2D(st, gd["dx"], gd["dxm"], gd["dy"], gd["dym"], pp["qav"], pp["gm"], pp["safety"],
                        margin=margin)
    run.reset(dtmin=pp["dtmin"])
    run.steps(nsteps)
    cur, info = run.state()
    assert info["nexec"] == nsteps and info["merr"] == 0
    assert info["dt"] == dts[-1]
    for k, tt in zip(F2, cur):
        assert_bitwise(to_host(tt), ref[k], k)


def sum_sequential(xs):
    t = 0.0
    for x in xs:
        t = t + x
    return t


def test_mlw2d_resident_stops_at_tend(dev, oracle):
    """t < tend loop test: steps after the end are no-ops."""
    from utrip_amd import cans
    pb = P.orszag_tang_2d(n=40, margin=1)
    ref, dts = _loop_2d_oracle(oracle, pb, 6)
    tend = sum_sequential(dts[:5]) * (1 + 1e-12)   # reached after step 6
    g, pp = pb.grid, pb.params
    st = [to_dev(pb.fields[k], dev) for k in F2]
    gd = {k: to_dev(v, dev) for k, v in g.items()}
    run = cans.MlwRun2D(st, gd["dx"], gd["dxm"], gd["dy"], gd["dym"], pp["qav"], pp["gm"], pp["safety"])
    run.reset(tend=tend, dtmin=pp["dtmin"])
    run.steps(11)
    cur, info = run.state()
    assert info["nexec"] == 6 and info["ns"] == 6
    for k, tt in zip(F2, cur):
        assert_bitwise(to_host(tt), ref[k], k)


def test_mlw2d_resident_dtmin_stop(dev):
    """dt < dtmin sets merr = 9001 on the device and freezes the state."""
    from utrip_amd import cans
    pb = P.orszag_tang_2d(n=40, margin=1)
    g, pp = pb.grid, pb.params
    st = [to_dev(pb.fields[k], dev) for k in F2]
    orig = [s.clone() for s in st]
    gd = {k: to_dev(v, dev) for k, v in g.items()}
    run = cans.MlwRun2D(st, gd["dx"], gd["dxm"], gd["dy"], gd["dym"], pp["qav"], pp["gm"], pp["safety"])
    run.reset(dtmin=1.0)
    run.steps(4)
    cur, info = run.state()
    assert info["merr"] == 9001 and info["nexec"] == 0 and info["ns"] == 1
    for a, b in zip(cur, orig):
        assert_bitwise(to_host(a), to_host(b))


@pytest.mark.parametrize("dim", [1, 2])
def test_cfl_merr_and_nan(dev, oracle, dim):
    """cfl_m: merr 9001 when dt < dtmin; a NaN cell is skipped by the min
    (cfl_m.f: `if (dtq < dt)` is false for NaN)."""
    from utrip_amd import cans
    if dim == 1:
        s = smooth_state_1d(200, seed=3)
        _, dx, _ = stretched_grid(200)
        s["pr"][57] = np.nan
        dto, merro, _ = oracle.cfl_m_1d(0.4, 1e-10, 1.4, s["bx"], s["ro"], s["pr"], s["vx"], s["vy"], s["by"], dx)
        t = {k: to_dev(v, dev) for k, v in s.items()}
        dtg, merrg = cans.cfl_m_1d(0.4, 1e-10, 1.4, t["bx"], t["ro"], t["pr"], t["vx"], t["vy"], t["by"],
                                   to_dev(dx, dev))
        assert dtg == dto and merrg == merro == 0
        dtg, merrg = cans.cfl_m_1d(0.4, 1.0, 1.4, t["bx"], t["ro"], t["pr"], t["vx"], t["vy"], t["by"],
                                   to_dev(dx, dev))
        assert merrg == 9001 and dtg == dto
    else:
        s = smooth_state_2d(50, 37, seed=4)
        _, dx, _ = stretched_grid(50)
        _, dy, _ = stretched_grid(37, seed=5)
        s["ro"][10, 20] = np.nan
        dto, merro, _ = oracle.cfl_m_2d(0.4, 1e-10, 1.4, s["ro"], s["pr"], s["vx"], s["vy"], s["bx"], s["by"],
                                        dx, dy)
        t = {k: to_dev(v, dev) for k, v in s.items()}
        dtg, merrg = cans.cfl_m_2d(0.4, 1e-10, 1.4, t["ro"], t["pr"], t["vx"], t["vy"], t["bx"], t["by"],
                                   to_dev(dx, dev), to_dev(dy, dev))
        assert dtg == dto and merrg == merro == 0


@pytest.mark.parametrize("kind", range(8))
def test_bnd_kinds(dev, oracle, kind):
    from utrip_amd import cans
    rng = np.random.default_rng(kind)
    for margin in (1, 2, 3):
        q = np.asfortranarray(rng.standard_normal(23))
        _, _, dxm = stretched_grid(23)
        sides = [(margin, margin)] if kind == 7 else [(0, margin), (1, margin)]
        for mbnd, mg in sides:
            r = q.copy()
            oracle.bnd_1d(kind, mbnd, mg, r, 0.37, dxm)
            t = to_dev(q, dev)
            cans.bnd_1d(kind, mbnd, mg, t, 0.37, to_dev(dxm, dev))
            assert_bitwise(to_host(t), r, f"1d kind {kind} side {mbnd} margin {margin}")
        q2 = np.asfortranarray(rng.standard_normal((13, 11)))
        for d in (0, 1):
            _, _, dm = stretched_grid(13 if d == 0 else 11)
            for mbnd, mg in sides:
                r = q2.copy(order="F")
                oracle.bnd_2d(d, kind, mbnd, mg, r, 0.37, dm)
                t = to_dev(q2, dev)
                cans.bnd_2d(d, kind, mbnd, mg, t, 0.37, to_dev(dm, dev))
                assert_bitwise(to_host(t), r, f"2d dir {d} kind {kind} side {mbnd} margin {margin}")
