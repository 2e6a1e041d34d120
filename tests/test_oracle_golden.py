"""The C restatement (oracle/) reproduces the golden fixtures written by the
reference Fortran itself (tests/golden/gen_golden.py), bit for bit: the
CANS main loop {cfl_m; engine; bnd} on small instances of every benchmark
problem and engine."""
from pathlib import Path

import numpy as np
import pytest

from tests.helpers import assert_bitwise

GOLDEN = Path(__file__).resolve().parent / "golden"
# engine fixtures (gen_golden.py); the tube set and the shipped spicule run
# have their own tests
CASES = sorted(p.stem for p in GOLDEN.glob("*.npz") if not p.stem.startswith(("tubes_", "spicule_")))

F = {1: ("ro", "pr", "vx", "vy", "by"), 2: ("ro", "pr", "vx", "vy", "bx", "by", "az"),
     3: ("ro", "pr", "vx", "vy", "vz", "bx", "by", "bz")}


def test_fixtures_present():
    assert len(CASES) >= 7, CASES


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_fixture(oracle, name):
    z = np.load(GOLDEN / f"{name}.npz")      # allow_pickle stays False
    dim = int(name[name.index("d_") - 1])
    engine = name[: name.index(str(dim))]
    gm, safety, qav, dtmin, margin, nsteps = z["params"]
    margin, nsteps = int(margin), int(nsteps)
    f = {k[3:]: np.asfortranarray(z[k]).copy(order="F") for k in z.files if k.startswith("in_")}
    g = {k[5:]: np.ascontiguousarray(z[k]) for k in z.files if k.startswith("grid_")}
    periodic = "orszag" in name or "blast" in name
    names = F[dim]
    dts = []
    for _ in range(nsteps):
        if dim == 1:
            dt, merr, _ = oracle.cfl_m_1d(safety, dtmin, gm, f["bx"], f["ro"], f["pr"], f["vx"], f["vy"], f["by"], g["dx"])
        elif dim == 2:
            dt, merr, _ = oracle.cfl_m_2d(safety, dtmin, gm, *[f[k] for k in names[:6]], g["dx"], g["dy"])
        else:
            dt, merr, _ = oracle.cfl_m_3d(safety, dtmin, gm, *[f[k] for k in names], g["dx"], g["dy"], g["dz"])
        assert merr == 0
        if engine == "mlw":
            if dim == 1:
                oracle.mlw_m_1d(*[f[k] for k in names], f["bx"], f["bxm"], dt, qav, gm, g["dx"], g["dxm"])
            elif dim == 2:
                oracle.mlw_m_2d(*[f[k] for k in names], dt, qav, gm, g["dx"], g["dxm"], g["dy"], g["dym"])
            else:
                oracle.mlw_m_3d(*[f[k] for k in names], dt, qav, gm, g["dx"], g["dxm"], g["dy"], g["dym"], g["dz"],
                                g["dzm"])
        elif engine == "roe":
            if dim == 1:
                oracle.roe_m_1d(*[f[k] for k in names], f["bx"], f["bxm"], dt, gm, g["dx"])
            elif dim == 2:
                oracle.roe_m_2d(*[f[k] for k in names], dt, gm, g["dx"], g["dy"])
            else:
                oracle.roe_m_3d(*[f[k] for k in names], dt, gm, g["dx"], g["dy"], g["dz"])
        else:
            oracle.hlld_engine(*[f[k] for k in names], gm, 1.0, dt, g["dx"], g["dy"], g["dz"], (1, 1, 1))
        for d in range(dim):
            for k in names:
                if periodic:
                    if dim == 2:
                        oracle.bnd_2d(d, 7, margin, margin, f[k])
                    else:
                        oracle.bnd_3d(d, 7, margin, margin, f[k])
                else:
                    oracle.bnd_1d(0, 0, margin, f[k])
                    oracle.bnd_1d(0, 1, margin, f[k])
        dts.append(dt)
    assert_bitwise(np.array(dts), z["dts"], "dt sequence")
    for k in names:
        assert_bitwise(f[k], z[f"out_{k}"], f"{name}:{k}")


def test_tube_set_fixture_regenerates():
    """The batched-tube workload fixture is the documented seeded draw."""
    from utrip_amd.problems import tube_set
    z = np.load(GOLDEN / "tubes_4096.npz")
    nx, beta, s = tube_set(4096, int(z["seed"]))
    assert int(z["seed"]) == 20160809
    assert np.array_equal(nx, z["nx"]) and np.array_equal(beta, z["beta"]) and np.array_equal(s, z["s"])
    assert nx.min() >= 256 and nx.max() <= 4096 and 8.5e6 < nx.sum() < 9.3e6
    assert np.allclose(2 * s / 1.5625, beta) and beta.min() >= 1e-2 and beta.max() <= 1e2
