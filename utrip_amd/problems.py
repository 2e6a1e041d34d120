"""Initial states of the benchmark problems, built as a CANS ``model`` +
``grdrdy`` + ``bnd`` prologue would build them (host side, numpy, FP64).

Arrays are Fortran-ordered with the reference's shapes ``(ix,)``, ``(ix,jx)``
or ``(ix,jx,kx)``, ghost layers included.  Magnetic fields are in the CANS
Gaussian units of the supported engines (pressure B^2/8pi), i.e. a field
given in Heaviside-Lorentz units is multiplied by sqrt(4 pi)
(cans1d/md_mhdshktb/hdmlw/model.f:45-55 does the same).

The same arrays feed the GPU path and the CPU checkers bit for bit.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict

import numpy as np

SQ4PI = math.sqrt(4.0 * math.pi)


def grdrdy(x: np.ndarray):
    """Metric arrays from cell positions (cans1d/common/grdrdy.f:28-42)."""
    ix = x.shape[0]
    dxm = np.empty(ix)
    dx = np.empty(ix)
    xm = np.empty(ix)
    dxm[: ix - 1] = x[1:] - x[:-1]
    dxm[ix - 1] = dxm[ix - 2]
    dx[1: ix - 1] = (dxm[: ix - 2] + dxm[1: ix - 1]) / 2.0
    dx[0] = dx[1]
    dx[ix - 1] = dx[ix - 2]
    xm[: ix - 1] = (x[: ix - 1] + x[1:]) / 2.0
    xm[ix - 1] = xm[ix - 2] + dx[ix - 2]
    return dx, xm, dxm


def periodic_fill(q: np.ndarray, margin: int, axes=None) -> None:
    """bdper{x,y,z} with margin0 = margin1 = margin, x first
    (cans2d/bc/bdperx.f:31-44, bdpery.f)."""
    axes = range(q.ndim) if axes is None else axes
    for ax in axes:
        n = q.shape[ax]
        src_lo = [slice(None)] * q.ndim
        dst_lo = [slice(None)] * q.ndim
        src_lo[ax] = slice(n - 2 * margin, n - margin)
        dst_lo[ax] = slice(0, margin)
        q[tuple(dst_lo)] = q[tuple(src_lo)]
        src_hi = [slice(None)] * q.ndim
        dst_hi = [slice(None)] * q.ndim
        src_hi[ax] = slice(margin, 2 * margin)
        dst_hi[ax] = slice(n - margin, n)
        q[tuple(dst_hi)] = q[tuple(src_hi)]


def bbtoaa(bx: np.ndarray, by: np.ndarray, dxm: np.ndarray, dym: np.ndarray, margin: int) -> np.ndarray:
    """Vector potential Az from (Bx, By) by the trapezoid rule
    (cans2d/common/bbtoaa.f:28-49), in the reference's order of operations."""
    ix, jx = bx.shape
    az = np.zeros((ix, jx), order="F")
    i = margin  # 0-based index of the reference's i = margin+1
    az[i, margin] = 0.0
    for j in range(margin + 1, jx):
        az[i, j] = az[i, j - 1] + 0.5 * (bx[i, j - 1] + bx[i, j]) * dym[j - 1]
    for j in range(margin - 1, -1, -1):
        az[i, j] = az[i, j + 1] - 0.5 * (bx[i, j] + bx[i, j + 1]) * dym[j]
    for ii in range(margin + 1, ix):
        az[ii, :] = az[ii - 1, :] - 0.5 * (by[ii - 1, :] + by[ii, :]) * dxm[ii - 1]
    for ii in range(margin - 1, -1, -1):
        az[ii, :] = az[ii + 1, :] + 0.5 * (by[ii, :] + by[ii + 1, :]) * dxm[ii]
    return az


@dataclass
class Problem:
    name: str
    fields: Dict[str, np.ndarray]
    grid: Dict[str, np.ndarray]
    params: Dict[str, float] = field(default_factory=dict)

    def copy(self) -> "Problem":
        return Problem(self.name, {k: v.copy(order="F") for k, v in self.fields.items()},
                       {k: v.copy() for k, v in self.grid.items()}, dict(self.params))


def _centres(n: int, margin: int, lo: float, hi: float):
    d0 = (hi - lo) / n
    x = lo + (np.arange(1, n + 2 * margin + 1) - margin - 0.5) * d0
    return x


def shock_tube_1d(nx: int = 1024, margin: int = 2, gm: float = 2.0, bx_hl: float = 0.75,
                  rho_l=1.0, p_l=1.0, by_l=1.0, rho_r=0.125, p_r=0.1, by_r=-1.0, x0=0.5) -> Problem:
    """Brio-Wu MHD shock tube on [0,1] with the jump at x0
    (the layout of cans1d/md_mhdshktb/hdmlw/model.f:16-55 shifted to [0,1]).
    Free boundaries (bdfrex) on all fields at both ends."""
    ix = nx + 2 * margin
    x = _centres(nx, margin, 0.0, 1.0)
    dx, xm, dxm = grdrdy(x)
    left = x <= x0
    ro = np.where(left, rho_l, rho_r).astype(np.float64)
    pr = np.where(left, p_l, p_r).astype(np.float64)
    by = np.where(left, by_l * SQ4PI, by_r * SQ4PI).astype(np.float64)
    vx = np.zeros(ix)
    vy = np.zeros(ix)
    bx = np.full(ix, bx_hl * SQ4PI)
    bxm = np.full(ix, bx_hl * SQ4PI)
    f = {k: np.asfortranarray(v) for k, v in dict(ro=ro, pr=pr, vx=vx, vy=vy, by=by, bx=bx, bxm=bxm).items()}
    for k in ("ro", "pr", "vx", "vy", "by"):
        q = f[k]
        q[:margin] = q[margin]
        q[ix - margin:] = q[ix - margin - 1]
    return Problem("shock_tube_1d", f, dict(x=x, dx=dx, dxm=dxm, xm=xm),
                   dict(gm=gm, margin=margin, safety=0.4, qav=3.0, tend=0.1, dtmin=1e-10))


def orszag_tang_2d(n: int = 1024, margin: int = 1, gm: float = 5.0 / 3.0) -> Problem:
    """Orszag-Tang vortex on the periodic box [0,2pi)^2:
    rho = 25/(36 pi), p = 5/(12 pi), v = (-sin y, sin x),
    B = (-sin y, sin 2x) (Gaussian; = HL field (-sin y, sin 2x)/sqrt(4 pi)
    times sqrt(4 pi)).  Az from bbtoaa, ghosts periodic."""
    two_pi = 2.0 * math.pi
    x = _centres(n, margin, 0.0, two_pi)
    y = _centres(n, margin, 0.0, two_pi)
    dx, xm, dxm = grdrdy(x)
    dy, ym, dym = grdrdy(y)
    X, Y = np.meshgrid(x, y, indexing="ij")
    ix, jx = X.shape
    ro = np.full((ix, jx), 25.0 / (36.0 * math.pi), order="F")
    pr = np.full((ix, jx), 5.0 / (12.0 * math.pi), order="F")
    vx = np.asfortranarray(-np.sin(Y))
    vy = np.asfortranarray(np.sin(X))
    bx = np.asfortranarray(-np.sin(Y))
    by = np.asfortranarray(np.sin(2.0 * X))
    f = dict(ro=ro, pr=pr, vx=vx, vy=vy, bx=bx, by=by)
    for q in f.values():
        periodic_fill(q, margin)
    az = bbtoaa(bx, by, dxm, dym, margin)
    periodic_fill(az, margin)
    f["az"] = az
    return Problem("orszag_tang_2d", f, dict(x=x, dx=dx, dxm=dxm, y=y, dy=dy, dym=dym),
                   dict(gm=gm, margin=margin, safety=0.4, qav=3.0, dtmin=1e-10))


def orszag_tang_box(nx: int, ny: int, margin: int = 1, pscale: float = 1.0, gm: float = 5.0 / 3.0) -> Problem:
    """The Orszag-Tang vortex of orszag_tang_2d on an nx x ny grid of the same
    periodic box, with the pressure scaled by pscale (the plasma beta with
    it).  nx = ny and pscale = 1 give orszag_tang_2d's arrays bit for bit."""
    two_pi = 2.0 * math.pi
    x = _centres(nx, margin, 0.0, two_pi)
    y = _centres(ny, margin, 0.0, two_pi)
    dx, xm, dxm = grdrdy(x)
    dy, ym, dym = grdrdy(y)
    X, Y = np.meshgrid(x, y, indexing="ij")
    ix, jx = X.shape
    ro = np.full((ix, jx), 25.0 / (36.0 * math.pi), order="F")
    pr = np.full((ix, jx), 5.0 / (12.0 * math.pi) * pscale, order="F")
    f = dict(ro=ro, pr=pr, vx=np.asfortranarray(-np.sin(Y)), vy=np.asfortranarray(np.sin(X)),
             bx=np.asfortranarray(-np.sin(Y)), by=np.asfortranarray(np.sin(2.0 * X)))
    for q in f.values():
        periodic_fill(q, margin)
    az = bbtoaa(f["bx"], f["by"], dxm, dym, margin)
    periodic_fill(az, margin)
    f["az"] = az
    return Problem("orszag_tang_box", f, dict(x=x, dx=dx, dxm=dxm, y=y, dy=dy, dym=dym),
                   dict(gm=gm, margin=margin, safety=0.4, qav=3.0, dtmin=1e-10))


def ot_run_set(nrun: int = 1024, seed: int = 20160811, n_lo: int = 32, n_hi: int = 256):
    """A seeded set of uneven periodic 2D runs: nx, ny uniform in
    [n_lo, n_hi], pressure scale log-uniform in [0.1, 10]."""
    rng = np.random.default_rng(seed)
    nx = rng.integers(n_lo, n_hi + 1, size=nrun)
    ny = rng.integers(n_lo, n_hi + 1, size=nrun)
    ps = 10.0 ** rng.uniform(-1.0, 1.0, size=nrun)
    return nx, ny, ps


def orszag_tang_25d(n: int = 1024, margin: int = 1, gm: float = 5.0 / 3.0) -> Problem:
    """The Orszag-Tang vortex with out-of-plane components for mlw_m3:
    as orszag_tang_2d plus vz = 0.2 sin(x + y), bz = 0.3 cos(x - y)
    (Gaussian units), ghosts periodic."""
    pb = orszag_tang_2d(n, margin, gm)
    X, Y = np.meshgrid(pb.grid["x"], pb.grid["y"], indexing="ij")
    vz = np.asfortranarray(0.2 * np.sin(X + Y))
    bz = np.asfortranarray(0.3 * np.cos(X - Y))
    for q in (vz, bz):
        periodic_fill(q, margin)
    f = {k: pb.fields[k] for k in ("ro", "pr", "vx", "vy")}
    f.update(vz=vz, bx=pb.fields["bx"], by=pb.fields["by"], bz=bz, az=pb.fields["az"])
    return Problem("orszag_tang_25d", f, pb.grid, pb.params)


def blast_3d(n: int = 384, margin: int = 2, gm: float = 5.0 / 3.0, gaussian_b: bool = True,
             p0=0.1, p1=10.0, r0=0.1) -> Problem:
    """MHD blast wave in the periodic unit cube: rho = 1, p = 0.1, an
    over-pressure sphere p = 10 of radius 0.1 at the centre (cell-centre
    test), B_HL = (1,1,0)/sqrt(2); times sqrt(4 pi) for the Gaussian engines,
    unchanged for the mu = 1 HLLD engine."""
    x = _centres(n, margin, 0.0, 1.0)
    dx, xm, dxm = grdrdy(x)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    shape = X.shape
    r2 = (X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2
    ro = np.ones(shape, order="F")
    pr = np.asfortranarray(np.where(r2 < r0 * r0, p1, p0))
    s = SQ4PI if gaussian_b else 1.0
    b0 = s / math.sqrt(2.0)
    f = dict(ro=ro, pr=pr, vx=np.zeros(shape, order="F"), vy=np.zeros(shape, order="F"),
             vz=np.zeros(shape, order="F"), bx=np.full(shape, b0, order="F"), by=np.full(shape, b0, order="F"),
             bz=np.zeros(shape, order="F"))
    for q in f.values():
        periodic_fill(q, margin)
    return Problem("blast_3d", f, dict(x=x, dx=dx, dxm=dxm, y=x.copy(), dy=dx.copy(), dym=dxm.copy(),
                                       z=x.copy(), dz=dx.copy(), dzm=dxm.copy()),
                   dict(gm=gm, margin=margin, safety=0.4, qav=3.0, dtmin=1e-10))


C3_SEED = 20160809


def tube_set(ntube: int = 4096, seed: int = C3_SEED, nx_lo: int = 256, nx_hi: int = 4096):
    """The batched-tube workload: ``ntube`` Brio-Wu shock tubes (gamma 2,
    Bx_HL 0.75, By_HL +-1, rho 1 / 0.125) with nx = rng.integers(nx_lo,
    nx_hi+1) and plasma beta of the left state log-uniform in [1e-2, 1e2]
    (beta = 2 p_L / |B_HL|^2 = 2 p_L / 1.5625), reached by scaling both
    pressures: p_L = s, p_R = 0.1 s with s = beta * 1.5625 / 2.
    rng = numpy.random.default_rng(seed).  Returns (nx, beta, s)."""
    rng = np.random.default_rng(seed)
    nx = rng.integers(nx_lo, nx_hi + 1, ntube)
    beta = 10.0 ** rng.uniform(-2.0, 2.0, ntube)
    s = beta * 1.5625 / 2.0
    return nx.astype(np.int64), beta, s


def tubes(nx, s, margin: int = 2):
    """Shock-tube problems (shock_tube_1d) for the (nx, pressure scale) pairs
    of tube_set, as the dicts Batch1D packs."""
    out = []
    for n, sc in zip(nx, s):
        pb = shock_tube_1d(nx=int(n), margin=margin, p_l=float(sc), p_r=0.1 * float(sc))
        d = {k: pb.fields[k] for k in ("ro", "pr", "vx", "vy", "by", "bx", "bxm")}
        d["dx"], d["dxm"] = pb.grid["dx"], pb.grid["dxm"]
        out.append(d)
    return out
