"""Checks of the C restatement that come from the mathematics rather than
from the reference: exact index maps of the boundary routines (the table of
cans1d/bc/README), closed forms of grdrdy and cfl_m, fixed points of uniform
states, conservation in a periodic box, and the mirror / transpose
symmetries of the equations.  A dropped term, a wrong sign or a swapped
index in the oracle fails at least one of them."""
import math

import numpy as np
import pytest

from tests.helpers import smooth_state_2d
from utrip_amd import problems as P

F1 = ("ro", "pr", "vx", "vy", "by")
F2 = ("ro", "pr", "vx", "vy", "bx", "by", "az")
F3 = ("ro", "pr", "vx", "vy", "vz", "bx", "by", "bz")


# ------------------------------------------------------------------ bnd index maps
# ix = 7, margin = 2: ghosts 1,2 and 6,7 (1-based), interior 3..5.
Q = np.array([10.0, 20.0, 30.0, 40.0, 50.0, 60.0, 70.0])
EXPECT = {   # (kind, side) -> array after the fill (README: qq(1)=qq(2) etc.)
    (0, 0): [30, 30, 30, 40, 50, 60, 70], (0, 1): [10, 20, 30, 40, 50, 50, 50],
    (1, 0): [40, 30, 30, 40, 50, 60, 70], (1, 1): [10, 20, 30, 40, 50, 50, 40],
    (2, 0): [-40, -30, 30, 40, 50, 60, 70], (2, 1): [10, 20, 30, 40, 50, -50, -40],
    (3, 0): [50, 40, 30, 40, 50, 60, 70], (3, 1): [10, 20, 30, 40, 50, 40, 30],
    (4, 0): [-50, -40, 0, 40, 50, 60, 70], (4, 1): [10, 20, 30, 40, 0, -40, -30],
    (5, 0): [7, 7, 30, 40, 50, 60, 70], (5, 1): [10, 20, 30, 40, 50, 7, 7],
}


@pytest.mark.parametrize("key", sorted(EXPECT))
def test_bnd_index_maps(oracle, key):
    kind, side = key
    q = Q.copy()
    oracle.bnd_1d(kind, side, 2, q, 7.0)
    np.testing.assert_array_equal(q, np.array(EXPECT[key], dtype=float))


def test_bnd_periodic_and_extrapolation(oracle):
    q = Q.copy()
    oracle.bnd_1d(7, 2, 2, q)              # bdperx(2,2): ghosts are images of the far interior
    np.testing.assert_array_equal(q, [40, 50, 30, 40, 50, 30, 40])
    x = np.arange(7.0) * 0.5
    dxm = np.full(7, 0.5)
    q = 3.0 * x + 1.0
    q[[0, 1, 5, 6]] = -99.0
    oracle.bnd_1d(6, 0, 2, q, 0.0, dxm)      # bdfrdx: linear data is reproduced exactly
    oracle.bnd_1d(6, 1, 2, q, 0.0, dxm)
    np.testing.assert_allclose(q, 3.0 * x + 1.0, rtol=0, atol=1e-14)


def test_grdrdy_closed_form(oracle):
    h = 0.125                                 # exact binary spacing
    x = np.asfortranarray((np.arange(1, 21) - 2.5) * h)
    dx, xm, dxm = oracle.grdrdy_1d(x)
    np.testing.assert_array_equal(dxm, np.full(20, h))
    np.testing.assert_array_equal(dx, np.full(20, h))
    np.testing.assert_array_equal(xm[:-1], (x[:-1] + x[1:]) / 2)
    assert xm[-1] == xm[-2] + h
    for a, b in zip(P.grdrdy(x), (dx, xm, dxm)):
        np.testing.assert_array_equal(a, b)


def test_cfl_closed_form(oracle):
    gm, safety = 5 / 3, 0.4
    ro, pr, vx, vy, vz, bx, by, bz = 1.3, 0.7, 0.4, -0.2, 0.1, 0.9, -0.5, 0.3
    n = 9
    one = lambda v: np.full(n, v)
    dx = one(0.1)
    # 1D: |vx| + sqrt(cs2 + ca2), ca2 = (bx^2+by^2)/(4 pi ro)
    ca2 = (bx * bx + by * by) / (4 * math.pi * ro)
    cs2 = gm * pr / ro
    want = safety * 0.1 / (abs(vx) + math.sqrt(cs2 + ca2))
    dt, merr, imin = oracle.cfl_m_1d(safety, 1e-10, gm, one(bx), one(ro), one(pr), one(vx), one(vy), one(by), dx)
    assert merr == 0 and imin == 2 and dt == pytest.approx(want, rel=4e-16)
    # 2D: sqrt(v^2 + cs2 + ca2), min(dx, dy)
    s2 = {k: np.full((n, n), v, order="F") for k, v in dict(ro=ro, pr=pr, vx=vx, vy=vy, bx=bx, by=by).items()}
    want2 = safety * 0.05 / math.sqrt(vx * vx + vy * vy + cs2 + ca2)
    dt, merr, loc = oracle.cfl_m_2d(safety, 1e-10, gm, *[s2[k] for k in F2[:6]], dx, np.full(n, 0.05))
    assert loc == (1, 1) and dt == pytest.approx(want2, rel=4e-16)
    # 3D with the REAL*4 pi of cans3d/common/cfl_m.f
    s3 = {k: np.full((n, n, n), v, order="F") for k, v in
          dict(ro=ro, pr=pr, vx=vx, vy=vy, vz=vz, bx=bx, by=by, bz=bz).items()}
    pai = float(np.float32(math.pi))
    ca2f = (bx * bx + by * by + bz * bz) * (0.25 / pai) / ro
    want3 = safety * 0.05 / math.sqrt(vx * vx + vy * vy + vz * vz + cs2 + ca2f)
    dt, merr, _ = oracle.cfl_m_3d(safety, 1e-10, gm, *[s3[k] for k in F3], dx, np.full(n, 0.05), dx)
    assert dt == pytest.approx(want3, rel=4e-16)
    assert dt != pytest.approx(safety * 0.05 / math.sqrt(vx * vx + vy * vy + vz * vz + cs2 +
                                                         (bx * bx + by * by + bz * bz) / (4 * math.pi * ro)), rel=1e-12)
    # merr and a NaN cell that the min skips
    bad = one(pr)
    bad[4] = np.nan
    dt2, merr, _ = oracle.cfl_m_1d(safety, 1e-10, gm, one(bx), one(ro), bad, one(vx), one(vy), one(by), dx)
    assert dt2 == dt2 and merr == 0
    _, merr, _ = oracle.cfl_m_1d(safety, 1.0, gm, one(bx), one(ro), one(pr), one(vx), one(vy), one(by), dx)
    assert merr == 9001


# ------------------------------------------------------------------ uniform states are fixed points
def _uniform(dim, n, v=(1.3, 0.7, 0.4, -0.2, 0.1, 0.9, -0.5, 0.3)):
    names = dict(ro=v[0], pr=v[1], vx=v[2], vy=v[3], vz=v[4], bx=v[5], by=v[6], bz=v[7], az=0.0)
    shape = (n,) * dim
    keys = {1: F1 + ("bx", "bxm"), 2: F2, 3: F3}[dim]
    return {k: np.full(shape, names.get(k, v[5]), order="F") for k in keys}


@pytest.mark.parametrize("engine,dim", [("mlw", 1), ("mlw", 2), ("mlw", 3), ("roe", 1), ("roe", 2), ("roe", 3),
                                        ("hlld", 3)])
def test_uniform_state_fixed_point(oracle, engine, dim):
    n = 10
    s = _uniform(dim, n)
    s0 = {k: v.copy() for k, v in s.items()}
    h = np.full(n, 0.1)
    if engine == "mlw":
        if dim == 1:
            oracle.mlw_m_1d(*[s[k] for k in F1], s["bx"], s["bxm"], 0.01, 3.0, 5 / 3, h, h)
        elif dim == 2:
            oracle.mlw_m_2d(*[s[k] for k in F2], 0.01, 3.0, 5 / 3, h, h, h, h)
        else:
            oracle.mlw_m_3d(*[s[k] for k in F3], 0.01, 3.0, 5 / 3, h, h, h, h, h, h)
    elif engine == "roe":
        if dim == 1:
            oracle.roe_m_1d(*[s[k] for k in F1], s["bx"], s["bxm"], 0.01, 5 / 3, h)
        elif dim == 2:
            oracle.roe_m_2d(*[s[k] for k in F2], 0.01, 5 / 3, h, h)
        else:
            oracle.roe_m_3d(*[s[k] for k in F3], 0.01, 5 / 3, h, h, h)
    else:
        oracle.hlld_engine(*[s[k] for k in F3], 5 / 3, 1.0, 0.01, h, h, h, (1, 1, 1))
    lo = 3 if engine != "mlw" else 1
    sl = tuple(slice(lo, n - lo) for _ in range(dim))
    for k in {1: F1, 2: F2, 3: F3}[dim]:
        if k == "az":   # Az evolves by -Ez dt, a constant source
            continue
        np.testing.assert_allclose(s[k][sl], s0[k][sl], rtol=4e-15, atol=1e-15, err_msg=k)


# ------------------------------------------------------------------ conservation in a periodic box
def _energy(f, gm, dim):
    v2 = f["vx"] ** 2 + f["vy"] ** 2 + (f["vz"] ** 2 if dim == 3 else 0)
    b2 = f["bx"] ** 2 + f["by"] ** 2 + (f["bz"] ** 2 if dim == 3 else 0)
    return f["pr"] / (gm - 1) + 0.5 * f["ro"] * v2 + b2 / (8 * math.pi)


@pytest.mark.parametrize("engine", ["mlw", "roe"])
def test_periodic_conservation_2d(oracle, engine):
    margin = 1 if engine == "mlw" else 3
    pb = P.orszag_tang_2d(n=32, margin=margin)
    f = {k: v.copy(order="F") for k, v in pb.fields.items()}
    g, pp = pb.grid, pb.params
    sl = (slice(margin, -margin),) * 2
    m0 = f["ro"][sl].sum()
    e0 = _energy(f, pp["gm"], 2)[sl].sum()
    px0 = (f["ro"] * f["vx"])[sl].sum()
    for _ in range(10):
        dt, _, _ = oracle.cfl_m_2d(pp["safety"], pp["dtmin"], pp["gm"], *[f[k] for k in F2[:6]], g["dx"], g["dy"])
        if engine == "mlw":
            oracle.mlw_m_2d(*[f[k] for k in F2], dt, pp["qav"], pp["gm"], g["dx"], g["dxm"], g["dy"], g["dym"])
        else:
            oracle.roe_m_2d(*[f[k] for k in F2], dt, pp["gm"], g["dx"], g["dy"])
        for d in (0, 1):
            for k in F2:
                oracle.bnd_2d(d, 7, margin, margin, f[k])
    assert abs(f["ro"][sl].sum() - m0) < 1e-12 * abs(m0)
    assert abs(_energy(f, pp["gm"], 2)[sl].sum() - e0) < 1e-12 * abs(e0)
    assert abs((f["ro"] * f["vx"])[sl].sum() - px0) < 1e-11
    assert np.abs(f["ro"] - pb.fields["ro"]).max() > 1e-4   # the flow did evolve


# ------------------------------------------------------------------ symmetries
def test_mirror_symmetry_1d(oracle):
    """x -> -x maps (ro, pr, vx, vy, By) to (ro, pr, -vx, vy, -By) with Bx kept."""
    pb = P.shock_tube_1d(nx=128, margin=2)
    f = {k: v.copy() for k, v in pb.fields.items()}
    g, pp = pb.grid, pb.params
    m = {k: v[::-1].copy() for k, v in f.items()}
    m["vx"] = -m["vx"]
    m["by"] = -m["by"]
    dx, dxm = g["dx"], g["dxm"]
    for _ in range(30):
        dt, _, _ = oracle.cfl_m_1d(pp["safety"], 1e-10, pp["gm"], f["bx"], f["ro"], f["pr"], f["vx"], f["vy"], f["by"], dx)
        dtm, _, _ = oracle.cfl_m_1d(pp["safety"], 1e-10, pp["gm"], m["bx"], m["ro"], m["pr"], m["vx"], m["vy"], m["by"], dx)
        assert dtm == pytest.approx(dt, rel=1e-13)
        oracle.mlw_m_1d(*[f[k] for k in F1], f["bx"], f["bxm"], dt, pp["qav"], pp["gm"], dx, dxm)
        oracle.mlw_m_1d(*[m[k] for k in F1], m["bx"], m["bxm"], dt, pp["qav"], pp["gm"], dx, dxm)
        for k in F1:
            for side in (0, 1):
                oracle.bnd_1d(0, side, 2, f[k])
                oracle.bnd_1d(0, side, 2, m[k])
    for k, sgn in zip(F1, (1, 1, -1, 1, -1)):
        np.testing.assert_allclose(sgn * m[k][::-1], f[k], rtol=1e-11, atol=1e-11, err_msg=k)


@pytest.mark.parametrize("engine", ["mlw", "roe"])
def test_transpose_symmetry_2d(oracle, engine):
    """Swapping x and y maps (vx, vy, Bx, By, Az) to (vy, vx, By, Bx, -Az)."""
    ix = jx = 24
    s = smooth_state_2d(ix, jx, seed=3)
    h = np.full(ix, 0.1)
    t = {k: np.asfortranarray(v.T) for k, v in s.items()}
    t["vx"], t["vy"] = t["vy"].copy(order="F"), t["vx"].copy(order="F")
    t["bx"], t["by"] = t["by"].copy(order="F"), t["bx"].copy(order="F")
    t["az"] = -t["az"]
    for st in (s, t):
        if engine == "mlw":
            oracle.mlw_m_2d(*[st[k] for k in F2], 2e-3, 3.0, 5 / 3, h, h, h, h)
        else:
            oracle.roe_m_2d(*[st[k] for k in F2], 2e-3, 5 / 3, h, h)
    back = {"ro": t["ro"].T, "pr": t["pr"].T, "vx": t["vy"].T, "vy": t["vx"].T, "bx": t["by"].T, "by": t["bx"].T,
            "az": -t["az"].T}
    sl = (slice(3, -3),) * 2
    for k in F2:
        np.testing.assert_allclose(back[k][sl], s[k][sl], rtol=1e-12, atol=1e-13, err_msg=k)
