"""The workloads at their full sizes: the 1024^2 Orszag-Tang vortex of the
headline run, bit-identical to the C restatement of the reference for a few
steps, and the 384^3 blast wave (Roe and HLLD, FP64 and FP32-flux) through
properties that do not depend on a CPU run of that size: a periodic box
conserves the total mass to rounding, and the FP32-flux state stays within
the 1e-5 relative L1 density bound of the FP64 one."""
import numpy as np
import pytest
import torch

from tests.helpers import assert_bitwise, to_dev, to_host
from utrip_amd import problems as P

pytestmark = pytest.mark.gpu

F1 = ("ro", "pr", "vx", "vy", "by")
F2 = ("ro", "pr", "vx", "vy", "bx", "by", "az")
F3 = ("ro", "pr", "vx", "vy", "vz", "bx", "by", "bz")


def case(*a):
    print("CASE", *a, flush=True)


def _ot_oracle_loop(oracle, pb, nsteps, engine="mlw"):
    """cfl_m, engine, bdperx then bdpery of the oracle (the reference main.f
    step order) for nsteps steps on a copy of pb's state."""
    g, pp = pb.grid, pb.params
    m = pb.params.get("margin", 1)
    f = {k: v.copy(order="F") for k, v in pb.fields.items()}
    dts = []
    for _ in range(nsteps):
        dt, merr, _loc = oracle.cfl_m_2d(pp["safety"], pp["dtmin"], pp["gm"], f["ro"], f["pr"], f["vx"], f["vy"],
                                         f["bx"], f["by"], g["dx"], g["dy"])
        assert merr == 0
        if engine == "mlw":
            oracle.mlw_m_2d(*[f[k] for k in F2], dt, pp["qav"], pp["gm"], g["dx"], g["dxm"], g["dy"], g["dym"])
        else:
            oracle.roe_m_2d(*[f[k] for k in F2], dt, pp["gm"], g["dx"], g["dy"])
        for d in (0, 1):
            for k in F2:
                oracle.bnd_2d(d, 7, m, m, f[k])
        dts.append(dt)
    return f, dts


def test_ot1024_headline_run_matches_oracle(dev, oracle):
    """100 steps of the fused periodic loop at 1024^2 (the headline workload)
    against cfl_m, mlw_m, bdperx, bdpery of the oracle, every field bit for
    bit, checked after 6 and after 100 steps."""
    from utrip_amd import cans
    case("OT 1024^2 fused loop vs oracle, 100 steps")
    pb = P.orszag_tang_2d(n=1024, margin=1)
    g, pp = pb.grid, pb.params
    st = [to_dev(pb.fields[k], dev) for k in F2]
    gd = {k: to_dev(v, dev) for k, v in g.items()}
    run = cans.MlwRun2D(st, gd["dx"], gd["dxm"], gd["dy"], gd["dym"], pp["qav"], pp["gm"], pp["safety"], margin=1)
    run.reset(dtmin=pp["dtmin"])
    f = {k: v.copy(order="F") for k, v in pb.fields.items()}
    done = 0
    for upto in (6, 100):
        sub = P.Problem(pb.name, f, g, pp)
        f, dts = _ot_oracle_loop(oracle, sub, upto - done)
        run.steps(upto - done)
        done = upto
        cur, info = run.state()
        assert info["nexec"] == upto and info["merr"] == 0 and info["dt"] == dts[-1]
        for k, t in zip(F2, cur):
            assert_bitwise(to_host(t), f[k], f"{k} after {upto} steps")


def test_ot1024_roe_matches_oracle(dev, oracle):
    """5 steps of the 2D Roe-MUSCL loop (margin 3, periodic) at 1024^2 against
    the oracle's cfl_m, roe_m, bdperx, bdpery, bit for bit."""
    from utrip_amd import cans
    case("OT 1024^2 Roe loop vs oracle, 5 steps")
    margin = 3
    pb = P.orszag_tang_2d(n=1024, margin=margin)
    pb.params["margin"] = margin
    ref, dts = _ot_oracle_loop(oracle, pb, 5, engine="roe")
    st = [to_dev(pb.fields[k], dev) for k in F2]
    gd = {k: to_dev(v, dev) for k, v in pb.grid.items() if k in ("dx", "dxm", "dy", "dym")}
    pp = pb.params
    run = cans.Run(2, cans.ENGINE_ROE, st, gd, pp["gm"], pp["safety"], qav=pp["qav"], margin=margin, bc=7)
    run.reset(dtmin=pp["dtmin"])
    run.steps(5)
    cur, info = run.state()
    assert info["nexec"] == 5 and info["merr"] == 0 and info["dt"] == dts[-1]
    for k, t in zip(F2, cur):
        assert_bitwise(to_host(t), ref[k], k)


def test_ot1024_full_run_conserves_and_repeats(dev):
    """The headline run at its full length (2000 steps at 1024^2): the
    periodic box conserves mass, momentum, energy and magnetic flux (sums of
    the conserved variables over the interior) to rounding (1e-10 relative), and a second run
    from the same state gives the same bits (the dt reduction is
    order-independent)."""
    from utrip_amd import cans
    case("OT 1024^2, 2000 steps, conservation and repeatability")
    pb = P.orszag_tang_2d(n=1024, margin=1)
    g, pp = pb.grid, pb.params
    gd = {k: to_dev(v, dev) for k, v in g.items()}
    gm = pp["gm"]
    pi8i = 1.0 / np.pi / 8.0

    def sums(t):
        ro, pr, vx, vy, bx, by = (x[1:-1, 1:-1] for x in t[:6])
        ee = pr / (gm - 1.0) + 0.5 * ro * (vx * vx + vy * vy) + pi8i * (bx * bx + by * by)
        return [float(q.sum()) for q in (ro, ro * vx, ro * vy, ee, bx, by)]

    finals = []
    for rep in range(2):
        st = [to_dev(pb.fields[k], dev) for k in F2]
        s0 = sums(st)
        run = cans.MlwRun2D(st, gd["dx"], gd["dxm"], gd["dy"], gd["dym"], pp["qav"], pp["gm"], pp["safety"],
                            margin=1)
        run.reset(dtmin=pp["dtmin"])
        run.steps(2000)
        cur, info = run.state()
        assert info["nexec"] == 2000 and info["merr"] == 0
        assert all(bool(torch.isfinite(t).all()) for t in cur)
        s1 = sums(cur)
        scale = [abs(s0[0]), abs(s0[0]), abs(s0[0]), abs(s0[3]), float(cur[4][1:-1, 1:-1].abs().sum()),
                 float(cur[5][1:-1, 1:-1].abs().sum())]
        for name, a, b, sc in zip(("mass", "x-momentum", "y-momentum", "energy", "bx", "by"), s0, s1, scale):
            print(name, a, b, abs(b - a) / sc, flush=True)
            assert abs(b - a) / sc < 1e-10, name
        finals.append([t.clone() for t in cur])
        del run, cur
    for k, a, b in zip(F2, *finals):
        assert torch.equal(a, b), k


def _c0_oracle(oracle, pb):
    """cans1d/md_mhdshktb/hdmlw/main.f:132-186 with the c0 parameters: ns=ns+1;
    cfl_m; t=t+dt; mlw_m; bdfrex at both ends; repeat while t < tend."""
    g, pp = pb.grid, pb.params
    m = pp["margin"]
    f = {k: v.copy() for k, v in pb.fields.items()}
    t, dts = 0.0, []
    while True:
        dt, merr, _ = oracle.cfl_m_1d(pp["safety"], pp["dtmin"], pp["gm"], f["bx"], f["ro"], f["pr"], f["vx"],
                                      f["vy"], f["by"], g["dx"])
        assert merr == 0
        t = t + dt
        oracle.mlw_m_1d(*[f[k] for k in F1], f["bx"], f["bxm"], dt, pp["qav"], pp["gm"], g["dx"], g["dxm"])
        for k in F1:
            oracle.bnd_1d(0, 0, m, f[k])
            oracle.bnd_1d(0, 1, m, f[k])
        dts.append(dt)
        if not t < pp["tend"]:
            return f, dts, t


def test_c0_shock_tube_to_tend_dropin(dev, oracle):
    """The c0 workload to its end (Brio-Wu, nx=1024, margin 2, to t=0.1, about
    a thousand steps) through the drop-in entry points, the host keeping t as
    main.f does: the same dt at every step and the same final state, bit for
    bit."""
    from utrip_amd import cans
    case("c0 shock tube to t=0.1, drop-in loop")
    pb = P.shock_tube_1d(nx=1024, margin=2)
    ref, dts, tref = _c0_oracle(oracle, pb)
    g, pp = pb.grid, pb.params
    t = {k: to_dev(v, dev) for k, v in pb.fields.items()}
    dx, dxm = to_dev(g["dx"], dev), to_dev(g["dxm"], dev)
    time_, gdts = 0.0, []
    while True:
        dt, merr = cans.cfl_m_1d(pp["safety"], pp["dtmin"], pp["gm"], t["bx"], t["ro"], t["pr"], t["vx"], t["vy"],
                                 t["by"], dx)
        assert merr == 0
        time_ = time_ + dt
        cans.mlw_m_1d(t["ro"], t["pr"], t["vx"], t["vy"], t["by"], t["bx"], t["bxm"], dt, pp["qav"], pp["gm"],
                      dx, dxm)
        for k in F1:
            cans.bnd_1d("fre", 0, 2, t[k])
            cans.bnd_1d("fre", 1, 2, t[k])
        gdts.append(dt)
        if not time_ < pp["tend"]:
            break
    print("steps", len(gdts), "t", time_, flush=True)
    assert len(gdts) == len(dts) and gdts == dts and time_ == tref
    for k in F1:
        assert_bitwise(to_host(t[k]), ref[k], k)


def test_c0_shock_tube_to_tend_resident(dev, oracle):
    """The same run on the device clock (cans_run, tend = 0.1): the clock stops
    the loop by itself at the step the reference stops at; t, the last dt and
    the final state equal the oracle's bit for bit."""
    from utrip_amd import cans
    case("c0 shock tube to t=0.1, resident run")
    pb = P.shock_tube_1d(nx=1024, margin=2)
    ref, dts, tref = _c0_oracle(oracle, pb)
    g, pp = pb.grid, pb.params
    st = [to_dev(pb.fields[k], dev) for k in F1]
    gd = {k: to_dev(g[k], dev) for k in ("dx", "dxm")}
    run = cans.Run(1, cans.ENGINE_MLW, st, gd, pp["gm"], pp["safety"], qav=pp["qav"], margin=2, bc=0,
                   bx1d=to_dev(pb.fields["bx"], dev), bxm1d=to_dev(pb.fields["bxm"], dev))
    run.reset(tend=pp["tend"], dtmin=pp["dtmin"])
    run.steps(len(dts) + 10)             # the clock ends the run at t >= tend
    cur, info = run.state()
    assert info["nexec"] == len(dts) and info["merr"] == 0, info
    assert info["t"] == tref and info["dt"] == dts[-1], info
    for k, x in zip(F1, cur):
        assert_bitwise(to_host(x), ref[k], k)


@pytest.mark.parametrize("engine", ["roe", "hlld"])
def test_blast384_conservation_and_fp32_flux(dev, engine):
    from utrip_amd import cans
    case("blast 384^3", engine)
    n, margin, nsteps = 384, 3, 3
    pb = P.blast_3d(n=n, margin=margin, gaussian_b=(engine != "hlld"))
    st0 = [torch.from_numpy(np.ascontiguousarray(pb.fields[k].T)).to(dev) for k in F3]
    gd = {k: torch.from_numpy(pb.grid[k]).to(dev) for k in ("dx", "dxm", "dy", "dym", "dz", "dzm")}
    del pb
    eng = cans.ENGINE_ROE if engine == "roe" else cans.ENGINE_HLLD
    sl = (slice(margin, -margin),) * 3                 # the periodic interior
    mass0 = float(st0[0][sl].sum())
    out = {}
    for mixed in (False, True):
        run = cans.Run(3, eng, [t.clone() for t in st0], gd, 5.0 / 3.0, 0.4, qav=3.0, mu=1.0, margin=margin,
                       bc=cans.BD_PER, mixed=mixed)
        run.reset(dtmin=1e-10)
        run.steps(nsteps)
        cur, info = run.state()
        assert info["nexec"] == nsteps and info["merr"] == 0, info
        out[mixed] = cur[0][sl].clone()
        assert all(bool(torch.isfinite(t).all()) for t in cur)
        del run, cur
        torch.cuda.empty_cache()
    drift = abs(float(out[False].sum()) - mass0) / mass0
    print("mass drift", drift, flush=True)
    assert drift < 1e-12
    rel = float((out[True] - out[False]).abs().sum() / out[False].abs().sum())
    print("rho rel L1 fp32-flux vs fp64", rel, flush=True)
    assert rel < 1e-5


def test_blast384_mlw_march_equals_three_kernels(dev):
    """3D MLW at the workload's size (384^3 blast, margin 2, 3 steps of the
    resident periodic run): the z-marching kernel (the default Cartesian
    path, 9 z-segments of 43 planes at this size) and the three kernels over
    HBM scratch, both held bitwise against the oracle at small sizes, give
    the same bits here on every field, and conserve mass."""
    from utrip_amd import cans
    case("blast 384^3 MLW, marching kernel vs three kernels")
    n, margin, nsteps = 384, 2, 3
    pb = P.blast_3d(n=n, margin=margin, gaussian_b=True)
    st0 = [torch.from_numpy(np.ascontiguousarray(pb.fields[k].T)).to(dev) for k in F3]
    gd = {k: torch.from_numpy(pb.grid[k]).to(dev) for k in ("dx", "dxm", "dy", "dym", "dz", "dzm")}
    del pb
    sl = (slice(margin, -margin),) * 3
    mass0 = float(st0[0][sl].sum())
    res = {}
    for path in (1, 0):
        cans.set_mlw3d_path(path)
        try:
            run = cans.Run(3, cans.ENGINE_MLW, [t.clone() for t in st0], gd, 5.0 / 3.0, 0.4, qav=3.0, margin=margin,
                           bc=cans.BD_PER)
            run.reset(dtmin=1e-10)
            run.steps(nsteps)
            cur, info = run.state()
        finally:
            cans.set_mlw3d_path(1)
        assert info["nexec"] == nsteps and info["merr"] == 0, info
        res[path] = ([t.clone() for t in cur], info["t"])
        del run, cur
        torch.cuda.empty_cache()
    assert res[1][1] == res[0][1]
    for k, a, b in zip(F3, res[1][0], res[0][0]):
        assert torch.equal(a, b), k
    drift = abs(float(res[1][0][0][sl].sum()) - mass0) / mass0
    print("mass drift", drift, flush=True)
    assert drift < 1e-12


def test_c0_shock_tube_to_tend_persistent(dev, oracle):
    """The same run as one persistent workgroup (the batched-tube kernel with
    one tube): same step count, t and final state as the oracle, bit for
    bit."""
    from utrip_amd import cans
    case("c0 shock tube to t=0.1, persistent kernel")
    pb = P.shock_tube_1d(nx=1024, margin=2)
    ref, dts, tref = _c0_oracle(oracle, pb)
    g, pp = pb.grid, pb.params
    tube = {k: pb.fields[k] for k in ("ro", "pr", "vx", "vy", "by", "bx", "bxm")}
    tube["dx"], tube["dxm"] = g["dx"], g["dxm"]
    b = cans.Batch1D([tube], pp["gm"], pp["qav"], pp["safety"], 2)
    b.run(tend=pp["tend"])
    torch.cuda.synchronize()
    assert int(b.ns[0]) == len(dts) and int(b.merr[0]) == 0 and float(b.t[0]) == tref
    got = b.tube(0)
    for k in F1:
        assert_bitwise(got[k], ref[k], k)
