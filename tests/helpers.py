"""Shared test helpers: reference-layout conversions and loop drivers."""
from __future__ import annotations

import ctypes as C

import numpy as np


def to_dev(a: np.ndarray, dev):
    """Fortran-ordered numpy (ix[,jx[,kx]]) -> contiguous torch tensor of
    shape (..., jx, ix) on `dev` (same memory layout as the reference)."""
    import torch
    return torch.from_numpy(np.ascontiguousarray(a.T)).to(dev)


def to_host(t) -> np.ndarray:
    """Inverse of to_dev."""
    return np.asfortranarray(t.detach().cpu().numpy().T)


def rel_l1(a: np.ndarray, b: np.ndarray) -> float:
    den = np.abs(b).sum()
    return float(np.abs(a - b).sum() / (den if den > 0 else 1.0))


def assert_bitwise(a: np.ndarray, b: np.ndarray, what: str = "") -> None:
    """Bit-identical, NaN-aware."""
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    ok = (a == b) | (np.isnan(a) & np.isnan(b))
    if not ok.all():
        idx = np.argwhere(~ok)
        first = tuple(idx[0])
        raise AssertionError(f"{what}: {len(idx)} mismatches, first at {first}: {a[first]!r} vs {b[first]!r}")


def smooth_state_1d(ix, seed=0):
    rng = np.random.default_rng(seed)
    x = np.linspace(0, 1, ix)
    k = rng.uniform(1, 4, 6)
    ph = rng.uniform(0, 6.28, 6)
    ro = 1.0 + 0.3 * np.sin(2 * np.pi * k[0] * x + ph[0])
    pr = 0.8 + 0.2 * np.cos(2 * np.pi * k[1] * x + ph[1])
    vx = 0.3 * np.sin(2 * np.pi * k[2] * x + ph[2])
    vy = 0.2 * np.cos(2 * np.pi * k[3] * x + ph[3])
    by = 1.5 * np.sin(2 * np.pi * k[4] * x + ph[4])
    bx = np.full(ix, 1.2)
    return {k_: np.asfortranarray(v) for k_, v in dict(ro=ro, pr=pr, vx=vx, vy=vy, by=by, bx=bx, bxm=bx.copy()).items()}


def stretched_grid(n, seed=1, lo=0.0, hi=1.0):
    """Non-uniform cell positions (exercises every metric coefficient)."""
    rng = np.random.default_rng(seed)
    w = 1.0 + 0.3 * rng.random(n)
    x = np.concatenate([[0.0], np.cumsum(w)])[:n]
    x = lo + (hi - lo) * x / x[-1]
    from utrip_amd.problems import grdrdy
    dx, xm, dxm = grdrdy(x)
    return x, dx, dxm


def smooth_state_2d(ix, jx, seed=0):
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid(np.linspace(0, 1, ix), np.linspace(0, 1, jx), indexing="ij")
    k = rng.uniform(1, 3, (7, 2))
    ph = rng.uniform(0, 6.28, 7)

    def w(n, amp, off):
        return np.asfortranarray(off + amp * np.sin(2 * np.pi * (k[n, 0] * X + k[n, 1] * Y) + ph[n]))
    return dict(ro=w(0, 0.3, 1.0), pr=w(1, 0.2, 0.8), vx=w(2, 0.3, 0.0), vy=w(3, 0.3, 0.0),
                bx=w(4, 1.0, 0.2), by=w(5, 1.0, -0.1), az=w(6, 0.5, 0.0))


def smooth_state_3d(ix, jx, kx, seed=0, bscale=1.0):
    rng = np.random.default_rng(seed)
    X, Y, Z = np.meshgrid(np.linspace(0, 1, ix), np.linspace(0, 1, jx), np.linspace(0, 1, kx), indexing="ij")

    def w(a, o):
        kk = rng.uniform(1, 2, 4)
        return np.asfortranarray(o + a * np.sin(2 * np.pi * (kk[0] * X + kk[1] * Y + kk[2] * Z) + kk[3]))
    return dict(ro=w(.3, 1), pr=w(.2, .8), vx=w(.3, 0), vy=w(.3, 0), vz=w(.3, 0), bx=w(bscale, .2 * bscale),
                by=w(bscale, -.1 * bscale), bz=w(bscale, .3 * bscale))


def grid3(ix, jx, kx):
    _, dx, dxm = stretched_grid(max(ix, 3), 1)
    _, dy, dym = stretched_grid(max(jx, 3), 4)
    _, dz, dzm = stretched_grid(max(kx, 3), 9)
    return dx[:ix].copy(), dxm[:ix].copy(), dy[:jx].copy(), dym[:jx].copy(), dz[:kx].copy(), dzm[:kx].copy()
