"""CPU checks behind the Roe adjoint: the generated reverse sweep of the face
flux is what utrip_amd/gen_roeflux_adj.py makes from the kernels' current
roeflux12, and the torch restatement of the Roe step (the reference of the
GPU adjoint tests) equals the oracle's loop cfl_m -> roe_m -> bdperx, bdpery
(cans2d/hdroe/roe_m.f) to rounding."""
import numpy as np
import torch

from utrip_amd import problems as P

F2 = ("ro", "pr", "vx", "vy", "bx", "by", "az")


def test_generated_roeflux_adjoint_is_current():
    from utrip_amd import gen_roeflux_adj as G
    assert G.OUT.read_text() == G.generate(), "rerun python3 -m utrip_amd.gen_roeflux_adj"


def test_torch_roe_step_matches_oracle(oracle):
    from tests import torch_roe2d as T
    M = 3
    pb = P.orszag_tang_2d(n=24, margin=M)
    pp, g = pb.params, pb.grid
    f = {k: v.copy(order="F") for k, v in pb.fields.items()}
    Pc = [torch.from_numpy(np.ascontiguousarray(pb.fields[k].T)) for k in F2[:6]]
    dx, dy = torch.from_numpy(g["dx"]), torch.from_numpy(g["dy"])
    for step in range(4):
        dt, merr, _ = oracle.cfl_m_2d(pp["safety"], pp["dtmin"], pp["gm"], f["ro"], f["pr"], f["vx"], f["vy"],
                                      f["bx"], f["by"], g["dx"], g["dy"])
        oracle.roe_m_2d(*[f[k] for k in F2], dt, pp["gm"], g["dx"], g["dy"])
        for d in (0, 1):
            for k in F2:
                oracle.bnd_2d(d, 7, M, M, f[k])
        Pc, dtt = T.step(Pc, dx, dy, pp["gm"], pp["safety"], M)
        assert abs(dt - float(dtt)) <= 1e-14 * dt
        for k, t in zip(F2[:6], Pc):
            ref = f[k].T
            assert float(np.abs(t.numpy() - ref).max() / np.abs(ref).max()) <= 1e-13, (step, k)
