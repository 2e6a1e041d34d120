"""Hand-written reverse mode of the periodic 2D Roe-MUSCL loop
(cans2d/hdroe/roe_m.f, roeflux_m.f, tvdminmod.f with bdperx/bdpery and
cfl_m; ghost width 3), behind the same torch.autograd.Function as the MLW
adjoint: L = sum_cells (rho_final - rho_target)^2 dx dy as a function of the
initial interior velocities and a uniform pressure scale beta.

Checked against (1) an FP64 CPU reference gradient: autograd of a torch
restatement of the step (tests/torch_roe2d.py) on CPU tensors, whose forward
is first checked against the engine; (2) fourth-order central differences of
the GPU loop itself.  The Roe flux switches (MAX of the eigenvalues with the
entropy-fix floor, the minmod limiter) make L piecewise smooth, so the FD
check is bounded relative to |g| as for the MLW loop with its viscosity
switch.  The Orszag-Tang state is exactly symmetric: its Roe averages of
the transverse field vanish exactly between mirror cells, where the flux's
square roots have no derivative (the kernel takes 0 there, autograd NaN).
The checks therefore start from a generic state: rho, p, bx, by of the
vortex times (1 + 1e-3 zeta) and v0 = v_OT (1 + 1e-3 eta), seeded."""
import numpy as np
import pytest
import torch

from utrip_amd import problems as P

pytestmark = pytest.mark.gpu
M = 3


def case(*a):
    print("CASE roe adjoint", *a, flush=True)


def setup(n, nsteps, dev, seed=7):
    from utrip_amd import adjoint
    pb = P.orszag_tang_2d(n=n, margin=M)
    names = ("ro", "pr", "vx", "vy", "bx", "by", "az")
    td = lambda a: torch.from_numpy(np.ascontiguousarray(a.T)).to(dev)
    rng = torch.Generator(device="cpu").manual_seed(seed)
    base = [td(pb.fields[k]) for k in names]
    for q in (0, 1, 4, 5):   # generic states: ro, p, bx, by times (1 + 1e-3 zeta), ghosts refilled
        zeta = torch.randn(base[q].shape, generator=rng, dtype=torch.float64).to(dev)
        base[q] = adjoint.periodic_fill(base[q] * (1 + 1e-3 * zeta), M)
    dx = torch.from_numpy(pb.grid["dx"]).to(dev)
    dy = torch.from_numpy(pb.grid["dy"]).to(dev)
    pp = pb.params
    solver = adjoint.PeriodicRoeAdjoint(dx, dy, pp["gm"], pp["safety"], M, nsteps)
    xi = lambda: torch.randn((n, n), generator=rng, dtype=torch.float64).to(dev)
    vx0 = base[2][M:-M, M:-M] * (1 + 1e-3 * xi())
    vy0 = base[3][M:-M, M:-M] * (1 + 1e-3 * xi())
    beta = torch.ones((), dtype=torch.float64, device=dev)
    with torch.no_grad():
        target = adjoint.FinalDensity.apply(vx0 * (1 + 0.01 * xi()), vy0 * (1 + 0.01 * xi()), beta, solver, base)
    w = float(pb.grid["dx"][M] * pb.grid["dy"][M])
    return solver, base, vx0, vy0, beta, target, w


def loss(vx0, vy0, beta, solver, base, target, w):
    from utrip_amd import adjoint
    rho = adjoint.FinalDensity.apply(vx0, vy0, beta, solver, base)
    return ((rho - target) ** 2).sum() * w


def test_torch_roe_step_matches_engine(dev):
    """The torch restatement of the Roe step (the CPU reference's forward)
    against the resident Roe engine: 4 steps of OT 24^2, every field and
    dt to 1e-12 relative."""
    from tests import torch_roe2d as T
    from utrip_amd import cans
    case("torch step vs engine")
    pb = P.orszag_tang_2d(n=24, margin=M)
    names = ("ro", "pr", "vx", "vy", "bx", "by", "az")
    td = lambda a: torch.from_numpy(np.ascontiguousarray(a.T)).to(dev)
    st = [td(pb.fields[k]) for k in names]
    gd = {k: td(pb.grid[k]) for k in ("dx", "dxm", "dy", "dym")}
    pp = pb.params
    run = cans.Run(2, cans.ENGINE_ROE, [t.clone() for t in st], gd, pp["gm"], pp["safety"], margin=M,
                   bc=cans.BD_PER)
    run.reset(dtmin=pp["dtmin"])
    Pc = [t.cpu() for t in st[:6]]
    dx, dy = gd["dx"].cpu(), gd["dy"].cpu()
    for k in range(4):
        run.steps(1)
        cur, info = run.state()
        Pc, dt = T.step(Pc, dx, dy, pp["gm"], pp["safety"], M)
        assert abs(float(dt) - info["dt"]) <= 1e-12 * info["dt"], (k, float(dt), info["dt"])
        for name, a, b in zip(names, Pc, cur):
            b = b.cpu()
            err = float((a - b).abs().max() / b.abs().max())
            assert err <= 1e-12, (k, name, err)


@pytest.mark.parametrize("n,nsteps", [(24, 6), (40, 12)])
def test_roe_adjoint_matches_cpu_reference_gradient(dev, n, nsteps):
    """The whole trajectory against an FP64 CPU reference gradient: autograd
    of the torch step on CPU tensors for the same steps, from the same
    initial velocities and pressure scale, with the same target; dL/dvx0,
    dL/dvy0, dL/dbeta to 1e-10 relative (max abs difference over max abs
    value)."""
    from tests import torch_roe2d as T
    case(f"OT {n}^2, {nsteps} steps vs CPU autograd of the torch step")
    solver, base, vx0, vy0, beta, target, w = setup(n, nsteps, dev)
    x = [vx0.clone().requires_grad_(True), vy0.clone().requires_grad_(True), beta.clone().requires_grad_(True)]
    L = loss(*x, solver, base, target, w)
    L.backward()
    g_gpu = [t.grad.detach().cpu() for t in x]
    cpu = torch.device("cpu")
    dx, dy = solver.dx.to(cpu), solver.dy.to(cpu)
    b = [t.to(cpu) for t in base]
    xc = [t.detach().to(cpu).clone().requires_grad_(True) for t in (vx0, vy0, beta)]

    def embed(full, inner):
        mid = torch.cat([full[M:-M, :M], inner, full[M:-M, -M:]], dim=1)
        return T.fill(torch.cat([full[:M], mid, full[-M:]], dim=0), M)

    Pc = [b[0], b[1] * xc[2], embed(b[2], xc[0]), embed(b[3], xc[1]), b[4], b[5]]
    for _ in range(nsteps):
        Pc, _dt = T.step(Pc, dx, dy, solver.gm, solver.safety, M)
    Lc = ((Pc[0][M:-M, M:-M] - target.to(cpu)) ** 2).sum() * w
    Lc.backward()
    print("L gpu", float(L), "L cpu", float(Lc), flush=True)
    assert abs(float(Lc) - float(L)) <= 1e-10 * abs(float(L))
    for name, a, c in zip(("vx0", "vy0", "beta"), g_gpu, (t.grad for t in xc)):
        r = float((a - c).abs().max() / c.abs().max())
        print(f"d L / d {name}: max rel diff {r:.2e}", flush=True)
        assert r <= 1e-10, (name, r)


def test_roe_adjoint_fd(dev):
    """OT 32^2, 4 steps (before the vortex steepens), 8 seeded random unit
    directions in (vx0, vy0, beta) and d = g/|g|: fourth-order central
    differences of the GPU loop at eps = 1e-6 against g.d, |FD - g.d| <=
    1e-6 |g| (measured 3e-10 .. 2.2e-7).  Later in the run the Roe step's
    switches (the minmod limiter of every face and field, the eigenvalue
    floors) are crossed by the +-eps runs: at 48^2 the FD error grows from
    ~1e-6 |g| after 8 steps to ~1e-3 |g| after 12 at any eps from 1e-5 to
    1e-7, while the gradient still equals the branch-exact CPU autograd
    reference to 1e-11 (test above) -- the difference quotient, not the
    gradient, degrades."""
    case("OT 32^2, 4 steps, FD")
    solver, base, vx0, vy0, beta, target, w = setup(32, 4, dev)
    x = [vx0.clone().requires_grad_(True), vy0.clone().requires_grad_(True), beta.clone().requires_grad_(True)]
    loss(*x, solver, base, target, w).backward()
    g = [t.grad.detach() for t in x]
    gn = float(torch.sqrt(sum((t * t).sum() for t in g)))
    rng = torch.Generator(device="cpu").manual_seed(11)

    def L_at(d, e):
        with torch.no_grad():
            return float(loss(vx0 + e * d[0], vy0 + e * d[1], beta + e * d[2], solver, base, target, w))

    dirs = []
    for _ in range(8):
        d = [torch.randn(vx0.shape, generator=rng, dtype=torch.float64).to(dev),
             torch.randn(vy0.shape, generator=rng, dtype=torch.float64).to(dev),
             torch.randn((), generator=rng, dtype=torch.float64).to(dev)]
        nrm = float(torch.sqrt(sum((t * t).sum() for t in d)))
        dirs.append([t / nrm for t in d])
    dirs.append([t / gn for t in g])
    eps = 1e-6
    for k, d in enumerate(dirs):
        fd = (L_at(d, -2 * eps) - 8 * L_at(d, -eps) + 8 * L_at(d, eps) - L_at(d, 2 * eps)) / (12 * eps)
        gd = float(sum((a * b).sum() for a, b in zip(g, d)))
        print(f"dir {k}: fd={fd:.12e} g.d={gd:.12e} rel(|g|)={abs(fd - gd) / gn:.2e}", flush=True)
        assert abs(fd - gd) <= 1e-6 * gn, (k, fd, gd, gn)


def test_roe_adjoint_recorded_half_equals_recomputed(dev):
    """The reverse sweep reading the recorded half-step primitives and the
    one recomputing them from the state give the same gradient bit for bit
    (the recompute runs the engine's own kernels on the same state), and the
    recorded final state equals the resident engine's run."""
    from utrip_amd import cans
    case("recorded vs recomputed half step")
    solver, base, vx0, vy0, beta, target, w = setup(24, 5, dev)
    x = [vx0.clone().requires_grad_(True), vy0.clone().requires_grad_(True), beta.clone().requires_grad_(True)]
    loss(*x, solver, base, target, w).backward()
    g1 = [t.grad.detach().clone() for t in x]
    lam = [torch.zeros_like(base[0]) for _ in range(6)]
    lam[0][M:-M, M:-M] = 1.0
    a = solver.backward(lam)
    b = solver.backward(lam, recompute_half=True)
    for p, q in zip(a, b):
        assert torch.equal(p, q)
    # the recording against the resident engine (same kernels, same order)
    grid = {"dx": solver.dx, "dxm": solver.dx, "dy": solver.dy, "dym": solver.dy}
    st = [t.clone() for t in solver._all[0]]
    run = cans.Run(2, cans.ENGINE_ROE, st, grid, solver.gm, solver.safety, margin=M, bc=cans.BD_PER)
    run.reset()
    run.steps(solver.nsteps)
    cur, info = run.state()
    assert info["dt"] == float(solver.dt_log[solver.nsteps - 1])
    for p, q in zip(cur, solver._all[solver.nsteps]):
        assert torch.equal(p, q)
    assert all(torch.isfinite(t).all() for t in g1)
