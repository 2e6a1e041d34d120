"""Batched periodic 2D runs (cans2d_batch_run): every run executes the
cans2d main loop {ns=ns+1; cfl_m; stop on merr; t=t+dt; mlw_m; bdperx;
bdpery; loop while ns<nstop and t<tend} (cans2d/md_mhdshktb/hdmlw/main.f:150-206)
inside one persistent kernel.  Each run must equal the oracle's loop bit for
bit (state, t, ns), whatever its size, pressure scale and ghost width, in any
launch shape, whether it stops on tend or on nstop."""
import numpy as np
import pytest
import torch

from tests.helpers import assert_bitwise
from utrip_amd import problems as P

pytestmark = pytest.mark.gpu

F2 = ("ro", "pr", "vx", "vy", "bx", "by", "az")


def make_run(nx, ny, m, pscale):
    pb = P.orszag_tang_box(nx, ny, margin=m, pscale=pscale)
    r = {k: v.copy(order="F") for k, v in pb.fields.items()}
    r.update({k: pb.grid[k] for k in ("dx", "dxm", "dy", "dym")})
    return r, pb.params


def oracle_run(oracle, r, pp, m, tend, nstop):
    """The main.f loop on the oracle: cfl_m, t=t+dt, mlw_m, bdperx, bdpery."""
    f = {k: r[k].copy(order="F") for k in F2}
    t, ns = 0.0, 0
    while True:
        ns += 1
        dt, merr, _ = oracle.cfl_m_2d(pp["safety"], pp["dtmin"], pp["gm"], *[f[k] for k in F2[:6]], r["dx"], r["dy"])
        assert merr == 0
        t = t + dt
        oracle.mlw_m_2d(*[f[k] for k in F2], dt, pp["qav"], pp["gm"], r["dx"], r["dxm"], r["dy"], r["dym"])
        for d in (0, 1):
            for k in F2:
                oracle.bnd_2d(d, 7, m, m, f[k])
        if not (ns < nstop and t < tend):
            return f, t, ns


def check_batch(oracle, specs, m, tend, nstop, shape, sample=None):
    from utrip_amd import cans
    runs, pp = [], None
    for nx, ny, ps in specs:
        r, pp = make_run(nx, ny, m, ps)
        runs.append(r)
    b = cans.Batch2D(runs, pp["gm"], pp["qav"], pp["safety"], m)
    b.run(tend=tend, nstop=nstop, block_threads=shape[0], blocks_per_cu=shape[1])
    torch.cuda.synchronize()
    t, ns, merr = b.t.cpu().numpy(), b.ns.cpu().numpy(), b.merr.cpu().numpy()
    ks = range(len(runs)) if sample is None else sample
    stops = set()
    for k in ks:
        r = runs[k]
        ref, tr, nsr = oracle_run(oracle, r, pp, m, tend, nstop)
        got = b.run_fields(k)
        for q in F2:
            assert_bitwise(got[q], ref[q], f"run {k} {q}")
        assert t[k] == tr and ns[k] == nsr and merr[k] == 0, (k, t[k], tr, ns[k], nsr)
        stops.add("nstop" if nsr == nstop else "tend")
    return stops


SPECS = [(10, 10, 0.5), (33, 20, 1.0), (64, 64, 2.0), (70, 45, 0.1), (130, 67, 10.0), (12, 200, 1.0), (5, 7, 1.0)]


@pytest.mark.parametrize("m,shape", [(1, (256, 2)), (1, (64, 8)), (2, (512, 1)), (3, (128, 2))])
def test_batch2d_matches_oracle_to_tend(dev, oracle, m, shape):
    print("CASE batch2d to tend, margin", m, "shape", shape, flush=True)
    stops = check_batch(oracle, SPECS, m, tend=0.05, nstop=10**6, shape=shape)
    assert stops == {"tend"}


def test_batch2d_matches_oracle_nstop(dev, oracle):
    print("CASE batch2d nstop and tend mixed", flush=True)
    stops = check_batch(oracle, SPECS, 1, tend=0.12, nstop=9, shape=(256, 2))
    assert stops == {"tend", "nstop"}


def test_batch2d_many_uneven_runs(dev, oracle):
    """300 runs of the seeded uneven set (nx, ny in 16..96, pressure scale
    0.1..10), every run checked."""
    print("CASE batch2d 300 uneven runs", flush=True)
    nx, ny, ps = P.ot_run_set(300, n_lo=16, n_hi=96)
    check_batch(oracle, list(zip(nx.tolist(), ny.tolist(), ps.tolist())), 1, tend=0.1, nstop=40, shape=(256, 2))


def test_batch2d_second_run_restarts_from_state(dev, oracle):
    """A second run() starts every run's loop again (t = 0, ns = 0) from the
    state the first left, reusing the workspace and its queue counter; it
    must equal the oracle's loop started from the first run's result."""
    from utrip_amd import cans
    print("CASE batch2d second run", flush=True)
    runs, pp = [], None
    for nx, ny, ps in SPECS[:5]:
        r, pp = make_run(nx, ny, 1, ps)
        runs.append(r)
    b = cans.Batch2D(runs, pp["gm"], pp["qav"], pp["safety"], 1)
    b.run(tend=0.1, nstop=3)
    b.run(tend=0.1, nstop=5)
    torch.cuda.synchronize()
    t, ns = b.t.cpu().numpy(), b.ns.cpu().numpy()
    for k, r in enumerate(runs):
        mid, _, _ = oracle_run(oracle, r, pp, 1, 0.1, 3)
        ref, tr, nsr = oracle_run(oracle, dict(r, **mid), pp, 1, 0.1, 5)
        got = b.run_fields(k)
        for q in F2:
            assert_bitwise(got[q], ref[q], f"run {k} {q}")
        assert t[k] == tr and ns[k] == nsr


def test_batch2d_tune_and_graph_replays(dev, oracle):
    """The autotuner leaves the state unchanged and picks a legal shape; the
    hipGraph of the batch run, replayed three times from the same initial
    state (with eager work and synchronisations between the replays), equals
    the oracle's loop for every run on every replay."""
    from utrip_amd import cans
    print("CASE batch2d tune + graph", flush=True)
    runs, pp = [], None
    for nx, ny, ps in SPECS:
        r, pp = make_run(nx, ny, 1, ps)
        runs.append(r)
    b = cans.Batch2D(runs, pp["gm"], pp["qav"], pp["safety"], 1)
    init = [t.clone() for t in b.f]
    bt, bpc, ms = b.tune(probe_steps=3)
    torch.cuda.synchronize()
    assert bt in (64, 128, 256, 512) and bpc >= 1 and bt * bpc <= 512
    for t0, t1 in zip(init, b.f):
        assert torch.equal(t0, t1)
    refs = [oracle_run(oracle, r, pp, 1, 0.05, 10**6) for r in runs]
    g = b.graph(tend=0.05, init=init)
    for rep in range(3):
        for t in b.f:            # eager work on the state between replays
            t.mul_(2.0)
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        t, ns = b.t.cpu().numpy(), b.ns.cpu().numpy()
        for k, (ref, tr, nsr) in enumerate(refs):
            got = b.run_fields(k)
            for q in F2:
                assert_bitwise(got[q], ref[q], f"replay {rep} run {k} {q}")
            assert t[k] == tr and ns[k] == nsr, (rep, k)
