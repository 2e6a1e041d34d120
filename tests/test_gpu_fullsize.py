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

F2 = ("ro", "pr", "vx", "vy", "bx", "by", "az")
F3 = ("ro", "pr", "vx", "vy", "vz", "bx", "by", "bz")


def case(*a):
    print("CASE", *a, flush=True)


def test_ot1024_headline_run_matches_oracle(dev, oracle):
    """6 steps of the fused periodic loop at 1024^2 against cfl_m, mlw_m,
    bdperx, bdpery of the oracle, every field bit for bit."""
    from utrip_amd import cans
    case("OT 1024^2 fused loop vs oracle, 6 steps")
    pb = P.orszag_tang_2d(n=1024, margin=1)
    g, pp = pb.grid, pb.params
    f = {k: v.copy(order="F") for k, v in pb.fields.items()}
    dts = []
    for _ in range(6):
        dt, merr, _loc = oracle.cfl_m_2d(pp["safety"], pp["dtmin"], pp["gm"], f["ro"], f["pr"], f["vx"], f["vy"],
                                         f["bx"], f["by"], g["dx"], g["dy"])
        assert merr == 0
        oracle.mlw_m_2d(*[f[k] for k in F2], dt, pp["qav"], pp["gm"], g["dx"], g["dxm"], g["dy"], g["dym"])
        for k in F2:
            oracle.bnd_2d(0, 7, 1, 1, f[k])
        for k in F2:
            oracle.bnd_2d(1, 7, 1, 1, f[k])
        dts.append(dt)
    st = [to_dev(pb.fields[k], dev) for k in F2]
    gd = {k: to_dev(v, dev) for k, v in g.items()}
    run = cans.MlwRun2D(st, gd["dx"], gd["dxm"], gd["dy"], gd["dym"], pp["qav"], pp["gm"], pp["safety"], margin=1)
    run.reset(dtmin=pp["dtmin"])
    run.steps(6)
    cur, info = run.state()
    assert info["nexec"] == 6 and info["merr"] == 0 and info["dt"] == dts[-1]
    for k, t in zip(F2, cur):
        assert_bitwise(to_host(t), f[k], k)


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
