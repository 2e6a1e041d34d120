"""Batched 1D tubes (cans1d_batch_run): every tube runs the cans1d main loop
{ns=ns+1; cfl_m; stop on merr; t=t+dt; mlw_m; bdfrex both ends; loop while
ns<nstop and t<tend} (cans1d/md_mhdshktb/hdmlw/main.f:132-186) inside one
persistent kernel.  FP64 must equal the oracle's loop per tube bit for bit
(state, t, ns), whatever the queue order and launch shape; the FP32-flux
variant stays within 1e-5 relative L1 in density."""
from pathlib import Path

import numpy as np
import pytest
import torch

from tests.helpers import assert_bitwise, rel_l1
from utrip_amd import problems as P

pytestmark = pytest.mark.gpu

F1 = ("ro", "pr", "vx", "vy", "by")
GM, QAV, SAFETY = 2.0, 3.0, 0.4


def case(*a):
    print("CASE batch1d", *a, flush=True)


def oracle_tube(oracle, d, margin, tend, nstop, dtmin=1e-10):
    f = {k: d[k].copy() for k in F1}
    t, ns, merr = 0.0, 0, 0
    while True:
        ns += 1
        dt, merr, _ = oracle.cfl_m_1d(SAFETY, dtmin, GM, d["bx"], *[f[k] for k in F1], d["dx"])
        if merr:
            break
        t = t + dt
        oracle.mlw_m_1d(*[f[k] for k in F1], d["bx"], d["bxm"], dt, QAV, GM, d["dx"], d["dxm"])
        for side in (0, 1):
            for k in F1:
                oracle.bnd_1d(0, side, margin, f[k])
        if not (ns < nstop and t < tend):
            break
    return f, t, ns, merr


def small_set(margin):
    nx = [7, 20, 63, 64, 130, 257]
    s = [0.0078, 0.1, 0.8, 2.5, 20.0, 78.0]
    return P.tubes(nx, s, margin=margin)


def make(tubes, margin, mixed=False):
    from utrip_amd import cans
    return cans.Batch1D(tubes, GM, QAV, SAFETY, margin, mixed=mixed)


@pytest.mark.parametrize("margin", [1, 2, 3])
def test_batch_matches_reference_loop(dev, oracle, margin):
    case("fp64 vs oracle, margin", margin)
    tubes = small_set(margin)
    b = make(tubes, margin)
    b.run(tend=0.02, nstop=60)
    torch.cuda.synchronize()
    t, ns, merr = b.t.cpu().numpy(), b.ns.cpu().numpy(), b.merr.cpu().numpy()
    for k, d in enumerate(tubes):
        ref, tr, nsr, mr = oracle_tube(oracle, d, margin, 0.02, 60)
        got = b.tube(k)
        for q in F1:
            assert_bitwise(got[q], ref[q], f"tube {k} {q}")
        assert t[k] == tr and ns[k] == nsr and merr[k] == mr, (k, t[k], tr, ns[k], nsr)


def test_batch_mixed_close(dev, oracle):
    case("fp32 flux vs fp64")
    tubes = small_set(2)
    b64, b32 = make(tubes, 2), make(tubes, 2, mixed=True)
    b64.run(tend=1e30, nstop=40)
    b32.run(tend=1e30, nstop=40)
    torch.cuda.synchronize()
    for k in range(len(tubes)):
        a, c = b64.tube(k), b32.tube(k)
        err = {q: rel_l1(c[q], a[q]) for q in ("ro", "pr", "by")}
        assert err["ro"] < 1e-5 and err["pr"] < 1e-4, (k, err)


def test_batch_shape_and_tuner_do_not_change_results(dev):
    case("launch shapes + tuner")
    tubes = small_set(2)
    ref = make(tubes, 2)
    ref.run(tend=0.02, nstop=60)
    torch.cuda.synchronize()
    want = [ref.tube(k) for k in range(len(tubes))]
    b = make(tubes, 2)
    before = [t.clone() for t in b.f]
    bt, bpc, ms = b.tune(probe_steps=5)
    torch.cuda.synchronize()
    assert bt in (64, 128, 256, 512) and 1 <= bpc <= 32
    for x, y in zip(b.f, before):
        assert torch.equal(x, y), "the tuner changed the state"
    for shape in ((64, 1), (512, 4), (bt, bpc)):
        c = make(tubes, 2)
        c.shape = shape
        c.run(tend=0.02, nstop=60)
        torch.cuda.synchronize()
        for k in range(len(tubes)):
            got = c.tube(k)
            for q in F1:
                assert_bitwise(got[q], want[k][q], f"shape {shape} tube {k} {q}")


def test_batch_graph_replays(dev):
    case("hipGraph replays")
    tubes = small_set(2)
    ref = make(tubes, 2)
    ref.run(tend=0.02, nstop=60)
    torch.cuda.synchronize()
    want_f = [t.clone() for t in ref.f]
    want_t, want_ns = ref.t.clone(), ref.ns.clone()
    b = make(tubes, 2)
    init = [t.clone() for t in b.f]
    g = b.graph(tend=0.02, nstop=60, init=init)
    for rep in range(3):
        junk = torch.randn(1000, device=dev).sum()          # eager work between replays
        g.launch()
        torch.cuda.synchronize()
        for x, y in zip(b.f, want_f):
            assert torch.equal(x, y), f"replay {rep}"
        assert torch.equal(b.t, want_t) and torch.equal(b.ns, want_ns), f"replay {rep}"
        del junk


def test_batch_dtmin_stop(dev):
    case("dtmin stop")
    tubes = small_set(2)
    from utrip_amd import cans
    b = cans.Batch1D(tubes, GM, QAV, SAFETY, 2, dtmin=1.0)
    before = [t.clone() for t in b.f]
    b.run(tend=1.0, nstop=10)
    torch.cuda.synchronize()
    assert (b.merr.cpu().numpy() == 9001).all() and (b.ns.cpu().numpy() == 1).all()
    assert (b.t.cpu().numpy() == 0.0).all()
    for x, y in zip(b.f, before):
        assert torch.equal(x, y)


def test_batch_rejects_bad_layout(dev):
    case("bad layout")
    from utrip_amd import cans
    tubes = small_set(2)
    b = make(tubes, 2)
    b.ix[0] = 4          # fewer than 2*margin+1 cells
    torch.cuda.synchronize()
    with pytest.raises(cans.CansError):
        b.run(tend=0.02, nstop=5)
    b.ix[0] = int(b.ix_host[0])
    b.order[0] = b.order[1]   # not a permutation
    torch.cuda.synchronize()
    with pytest.raises(cans.CansError):
        b.run(tend=0.02, nstop=5)


def test_batch_full_set_sample(dev, oracle):
    """The 4096-tube workload (seeded fixture) for 25 steps; the longest,
    the shortest and the extreme-beta tubes equal the oracle bit for bit."""
    case("4096 tubes, sampled vs oracle")
    fx = np.load(Path(__file__).resolve().parent / "golden" / "tubes_4096.npz")
    nx, beta, s = P.tube_set(4096, int(fx["seed"]))
    assert (nx == fx["nx"]).all() and np.array_equal(s, fx["s"])
    tubes = P.tubes(nx, s, margin=2)
    b = make(tubes, 2)
    b.run(tend=1e30, nstop=25)
    torch.cuda.synchronize()
    ns = b.ns.cpu().numpy()
    assert (ns == 25).all()
    for k in sorted({int(np.argmax(nx)), int(np.argmin(nx)), int(np.argmax(beta)), int(np.argmin(beta))}):
        ref, _, _, _ = oracle_tube(oracle, tubes[k], 2, 1e30, 25)
        got = b.tube(k)
        for q in F1:
            assert_bitwise(got[q], ref[q], f"tube {k} (nx={nx[k]}, beta={beta[k]:.3g}) {q}")


def test_full_tube_set_fp32_flux_bound(dev):
    """The whole batched workload (4096 seeded tubes, nx 256..4096, beta
    1e-2..1e2) to t = 0.1 in FP64, then with FP32 fluxes for exactly the
    steps each FP64 tube took: the density of every tube within 1e-5
    relative L1 of FP64 (the bound the mixed-precision forms are held to).
    Run to t = 0.1 on its own
    clock instead, a tube whose FP64 time lands within rounding of 0.1 can
    take one step more or less in FP32; the tubes that keep the FP64 step
    count meet the same bound."""
    from utrip_amd import cans
    case("full 4096-tube set, fp32 flux vs fp64")
    nx, beta, s = P.tube_set(4096)
    tubes = P.tubes(nx, s, margin=2)
    b64 = cans.Batch1D(tubes, GM, QAV, SAFETY, 2, tune_room=False)
    b64.run(tend=0.1)
    torch.cuda.synchronize()
    ns64 = b64.ns.clone()
    assert int(b64.merr.abs().max()) == 0 and bool((b64.t >= 0.1).all())
    ref = [t.clone() for t in b64.f]
    off, ixs = b64.off_host, b64.ix_host
    del b64

    def rel(f):
        r = np.empty(len(tubes))
        for k in range(len(tubes)):
            sl = slice(int(off[k]), int(off[k] + ixs[k]))
            r[k] = float((f[0][sl] - ref[0][sl]).abs().sum() / ref[0][sl].abs().sum())
        return r

    b32 = cans.Batch1D(tubes, GM, QAV, SAFETY, 2, mixed=True, tune_room=False)
    b32.run_steps(ns64)
    torch.cuda.synchronize()
    assert torch.equal(b32.ns, ns64) and int(b32.merr.abs().max()) == 0
    r = rel(b32.f)
    print("same steps: rho rel L1 max", r.max(), "median", np.median(r), "worst tube", int(r.argmax()), flush=True)
    assert r.max() <= 1e-5
    b32 = cans.Batch1D(tubes, GM, QAV, SAFETY, 2, mixed=True, tune_room=False)
    b32.run(tend=0.1)
    torch.cuda.synchronize()
    same = (b32.ns == ns64).cpu().numpy()
    r = rel(b32.f)
    print("to t=0.1: tubes with another step count", int((~same).sum()), "max rel L1 of the others",
          r[same].max(), flush=True)
    assert (~same).sum() <= 16
    assert r[same].max() <= 1e-5
