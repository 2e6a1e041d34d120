"""The helper routines of the hot path with their own reference argument
lists (the names the drop-in libraries export): mlwhalf, mlwfull, mlwartv,
mlwsrch, mlwsrcf in 1D/2D/3D, roeflux_m and tvdminmod in 1D/2D/3D,
hlld_flux and hlld_mh -- the oracle against the reference routines compiled
from their sources (oracle/_ref), bit for bit, on random inputs with every
output array pre-filled (the routines update du/dan in place and leave
entries they do not own untouched)."""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import assert_bitwise, grid3, stretched_grid


def rnd(rng, *shape, lo=-1.0, hi=1.0):
    return np.asfortranarray(rng.uniform(lo, hi, shape))


def pos(rng, *shape):
    return np.asfortranarray(rng.uniform(0.5, 1.5, shape))


def both(oracle, ref, oname, rname, args, outs, ref_args=None):
    """Run oracle.<oname>(*args) and ref.<rname>(*ref_args) on copies; compare
    the arrays at positions `outs`."""
    a = [x.copy(order="F") if isinstance(x, np.ndarray) else x for x in args]
    b = [x.copy(order="F") if isinstance(x, np.ndarray) else x for x in (ref_args or args)]
    oracle.call(oname, *a)
    ref.call(rname, *b)
    for k in outs:
        assert_bitwise(a[k], b[k], f"{rname} argument {k}")


@pytest.mark.parametrize("ix", [5, 40])
def test_mlw_helpers_1d(oracle, ref1d, ix):
    rng = np.random.default_rng(ix)
    _, dx, dxm = stretched_grid(ix, seed=3)
    dxi, dxim = 1.0 / dx, 1.0 / dxm
    u, f, q, s = rnd(rng, ix), rnd(rng, ix), pos(rng, ix), rnd(rng, ix)
    un, du = rnd(rng, ix), rnd(rng, ix)
    ux0, ux1 = pos(rng, ix), pos(rng, ix)
    dt = 1.3e-3
    both(oracle, ref1d, "or1d_mlwhalf", "mlwhalf", [u, un, du, dt, f, dxi, dxim, ix], [1, 2])
    both(oracle, ref1d, "or1d_mlwfull", "mlwfull", [du, dt, f, dxi, ix], [0])
    both(oracle, ref1d, "or1d_mlwartv", "mlwartv", [u, du, dt, q, dxi, dxim, ix], [1])


@pytest.mark.parametrize("ix", [6, 33])
def test_mlw_source_helpers_1d(oracle, refsp, ix):
    rng = np.random.default_rng(ix + 1)
    un, du, s = rnd(rng, ix), rnd(rng, ix), rnd(rng, ix)
    ux0, ux1 = pos(rng, ix), pos(rng, ix)
    both(oracle, refsp, "or1d_mlwsrch", "mlwsrch", [un, du, 2e-3, s, ix], [0, 1])
    both(oracle, refsp, "or1d_mlwsrcf", "mlwsrcf", [du, 2e-3, s, ux0, ux1, ix], [0])


@pytest.mark.parametrize("shape", [(3, 3), (17, 12)])
def test_mlw_helpers_2d(oracle, ref2d, shape):
    ix, jx = shape
    rng = np.random.default_rng(ix * jx)
    _, dx, dxm = stretched_grid(ix, seed=3)
    _, dy, dym = stretched_grid(jx, seed=5)
    dxi, dxim, dyi, dyim = 1 / dx, 1 / dxm, 1 / dy, 1 / dym
    u, fx, fy = rnd(rng, ix, jx), rnd(rng, ix, jx), rnd(rng, ix, jx)
    qx, qy, s = pos(rng, ix, jx), pos(rng, ix, jx), rnd(rng, ix, jx)
    un, du = rnd(rng, ix, jx), rnd(rng, ix, jx)
    ux0, ux1, uy0, uy1 = pos(rng, ix), pos(rng, ix), pos(rng, jx), pos(rng, jx)
    dt = 1.1e-3
    both(oracle, ref2d, "or2d_mlwhalf", "mlwhalf", [u, un, du, dt, fx, dxi, dxim, ix, fy, dyi, dyim, jx], [1, 2])
    both(oracle, ref2d, "or2d_mlwfull", "mlwfull", [du, dt, fx, dxi, ux0, ux1, ix, fy, dyi, uy0, uy1, jx], [0])
    both(oracle, ref2d, "or2d_mlwartv", "mlwartv", [u, du, dt, qx, dxi, dxim, ix, qy, dyi, dyim, jx], [1])
    both(oracle, ref2d, "or2d_mlwsrch", "mlwsrch", [un, du, dt, s, ix, jx], [0, 1])
    both(oracle, ref2d, "or2d_mlwsrcf", "mlwsrcf", [du, dt, s, ux0, ux1, ix, uy0, uy1, jx], [0])


@pytest.mark.parametrize("shape", [(3, 3, 3), (9, 7, 6)])
def test_mlw_helpers_3d(oracle, ref3d, shape):
    ix, jx, kx = shape
    rng = np.random.default_rng(ix * jx * kx)
    dx, dxm, dy, dym, dz, dzm = grid3(ix, jx, kx)
    m = [1 / dx, 1 / dxm, 1 / dy, 1 / dym, 1 / dz, 1 / dzm]
    sh = (ix, jx, kx)
    u, fx, fy, fz = rnd(rng, *sh), rnd(rng, *sh), rnd(rng, *sh), rnd(rng, *sh)
    qx, qy, qz, s = pos(rng, *sh), pos(rng, *sh), pos(rng, *sh), rnd(rng, *sh)
    un, du = rnd(rng, *sh), rnd(rng, *sh)
    w = [pos(rng, n) for n in (ix, ix, jx, jx, kx, kx)]
    dt = 0.9e-3
    both(oracle, ref3d, "or3d_mlwhalf", "mlwhalf",
         [u, un, du, dt, fx, m[0], m[1], ix, fy, m[2], m[3], jx, fz, m[4], m[5], kx], [1, 2])
    both(oracle, ref3d, "or3d_mlwfull", "mlwfull",
         [du, dt, fx, m[0], w[0], w[1], ix, fy, m[2], w[2], w[3], jx, fz, m[4], w[4], w[5], kx], [0])
    both(oracle, ref3d, "or3d_mlwartv", "mlwartv",
         [u, du, dt, qx, m[0], m[1], ix, qy, m[2], m[3], jx, qz, m[4], m[5], kx], [1])
    both(oracle, ref3d, "or3d_mlwsrch", "mlwsrch", [un, du, dt, s, ix, jx, kx], [0, 1])
    both(oracle, ref3d, "or3d_mlwsrcf", "mlwsrcf", [du, dt, s, w[0], w[1], ix, w[2], w[3], jx, w[4], w[5], kx], [0])


def _states(rng, shape, nf):
    """Left/right states (..., 2) in Fortran order: rho, p positive, the rest signed."""
    out = []
    for q in range(nf):
        a = pos(rng, *shape, 2) if q < 2 else rnd(rng, *shape, 2)
        out.append(a)
    return out


@pytest.mark.parametrize("ix", [2, 30])
def test_roeflux_tvdminmod_1d(oracle, ref1d, ix):
    rng = np.random.default_rng(100 + ix)
    w = _states(rng, (ix,), 6)
    fl = [rnd(rng, ix) for _ in range(5)]
    both(oracle, ref1d, "or1d_roeflux_m", "roeflux_m", [*fl, 5 / 3, *w, ix], range(5))
    da, daw = rnd(rng, ix), rnd(rng, ix, 2)
    both(oracle, ref1d, "or1d_tvdminmod", "tvdminmod", [da, daw, ix], [1])


@pytest.mark.parametrize("mdir", [1, 2])
def test_roeflux_tvdminmod_2d(oracle, ref2d, mdir):
    ix, jx = 11, 8
    rng = np.random.default_rng(200 + mdir)
    w = _states(rng, (ix, jx), 6)
    fl = [rnd(rng, ix, jx) for _ in range(5)]
    both(oracle, ref2d, "or2d_roeflux_m", "roeflux_m", [*fl, 5 / 3, *w, ix, jx], range(5))
    da, daw = rnd(rng, ix, jx), rnd(rng, ix, jx, 2)
    both(oracle, ref2d, "or2d_tvdminmod", "tvdminmod", [mdir, da, daw, ix, jx], [2])


@pytest.mark.parametrize("mdir", [1, 2, 3])
def test_roeflux_tvdminmod_3d(oracle, ref3d, mdir):
    ix, jx, kx = 7, 6, 5
    rng = np.random.default_rng(300 + mdir)
    w = _states(rng, (ix, jx, kx), 8)
    fl = [rnd(rng, ix, jx, kx) for _ in range(7)]
    both(oracle, ref3d, "or3d_roeflux_m", "roeflux_m", [*fl, 5 / 3, *w, ix, jx, kx], range(7))
    da, daw = rnd(rng, ix, jx, kx), rnd(rng, ix, jx, kx, 2)
    both(oracle, ref3d, "or3d_tvdminmod", "tvdminmod", [mdir, da, daw, ix, jx, kx], [2])


def test_hlld_flux(oracle, refhlld):
    ix, jx, kx = 6, 5, 4
    rng = np.random.default_rng(400)
    row, prw, vxw, vyw, vzw, byw, bzw = _states(rng, (ix, jx, kx), 7)
    bx = rnd(rng, ix, jx, kx)
    bx[0, 0, 0] = 0.0                      # the bxs == 0 branch (hlld_flux.f90:353-414)
    fl = [rnd(rng, ix, jx, kx) for _ in range(7)]
    args = [row, prw, vxw, vyw, vzw, bx, byw, bzw, 5 / 3, ix, jx, kx, 1e-10, *fl]
    both(oracle, refhlld, "orh_hlld_flux", "hlld_flux", args, range(13, 20))


@pytest.mark.parametrize("mstage", [1, 2])
@pytest.mark.parametrize("mf", [(1, 1, 1), (1, 0, 1), (0, 1, 0)])
def test_hlld_mh(oracle, refhlld, mf, mstage):
    """hlld_mh on the cells the reference defines: the updated conserved
    arrays on the interior 4..n-3 of every active direction and i < ix (its
    automatic arrays are unset at the last face, which reaches the ghost
    cells)."""
    ix, jx, kx = 10, 9, 8
    rng = np.random.default_rng(500 + mstage)
    sh = (ix, jx, kx)
    prim = [pos(rng, *sh), pos(rng, *sh)] + [rnd(rng, *sh) for _ in range(6)]
    un = [rnd(rng, *sh) for _ in range(8)]
    dx, _, dy, _, dz, _ = grid3(ix, jx, kx)
    mfd = np.array(mf, dtype=np.int32)
    a = [x.copy(order="F") for x in prim + un]
    b = [x.copy(order="F") for x in prim + un]
    oracle.call("orh_hlld_mh", *a, 5 / 3, 1.2e-3, dx, dy, dz, ix, jx, kx, mfd, mstage)
    refhlld.call("hlld_mh", *b, 5 / 3, 1.2e-3, dx, dy, dz, ix, jx, kx, mfd.copy(), mstage)
    # the y and z sweeps zero the normal-field flux on i = 1..ix-1 only
    # (hlld_mh.f:88-94, :103-109), so at i = ix it is the unset automatic array
    sl = tuple(slice(3, n - 3) if f else slice(0, n - 1 if d == 0 else n) for d, (n, f) in enumerate(zip(sh, mf)))
    for q in range(8, 16):
        assert_bitwise(a[q][sl], b[q][sl], f"hlld_mh conserved {q - 8}")
