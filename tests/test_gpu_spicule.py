"""The UTRIP spicule model on the GPU against the C restatement (oracle):
mlw_m_bg, trqalfven and chkdav as single calls, and the whole time loop
(md_spicule/hdmlw/main.f:169-227) for several realizations in one
persistent kernel, bit for bit in FP64."""
import numpy as np
import pytest
import torch

from tests.helpers import assert_bitwise
from tests.test_oracle_spicule import F, random_geometry, random_state
from utrip_amd import spicule

pytestmark = pytest.mark.gpu


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)


@pytest.mark.parametrize("ix", [4, 65, 300, 1026])
def test_mlw_m_bg_call(oracle, dev, ix):
    print(f"mlw_m_bg ix={ix}", flush=True)
    _, g = random_geometry(ix, ix)
    s = random_state(ix, ix)
    a = {k: v.copy() for k, v in s.items()}
    oracle.mlw_m_bg_1d(*[a[k] for k in F], g, 2.1e-3, 1.0, 5.0 / 3.0)
    d = {k: T(v, dev) for k, v in s.items()}
    spicule.mlw_m_bg(*[d[k] for k in F], {k: T(v, dev) for k, v in g.items()}, 2.1e-3, 1.0, 5.0 / 3.0)
    torch.cuda.synchronize()
    for k in F:
        assert_bitwise(d[k].cpu().numpy(), a[k], k)


def test_trqalfven_and_chkdav_calls(oracle, dev):
    sp = spicule.setup()
    ix = sp.ix
    rng = np.random.default_rng(3)
    vy = np.asfortranarray(0.01 * rng.standard_normal(ix))
    a = vy.copy()
    s_o = oracle.trqalfven(a, 2.5e-3, 0.0, np.asfortranarray(sp.x), np.asfortranarray(sp.geom["rr"]), 99991)
    d = T(vy, dev)
    s_g = spicule.trqalfven(d, 2.5e-3, T(spicule.torque_table(sp.x), dev), T(sp.geom["rr"], dev), 99991)
    torch.cuda.synchronize()
    assert s_g == s_o
    assert_bitwise(d.cpu().numpy(), a, "vy")
    da = np.asfortranarray(rng.uniform(-1e-9, 3e-9, ix))
    da[7] = np.nan
    vx = np.asfortranarray(rng.standard_normal(ix))
    da_o, vx_o = da.copy(), vx.copy()
    n_o = oracle.chkdav_1d(da_o, vx_o, 1e-9)
    dd, dv = T(da, dev), T(vx, dev)
    n_g = spicule.chkdav_dev(dd, dv, 1e-9)
    assert n_g == n_o > 0
    assert_bitwise(dd.cpu().numpy(), da_o, "da")
    assert_bitwise(dv.cpu().numpy(), vx_o, "vx")


def oracle_runs(oracle, sp, seeds, nstop, tend=spicule.TEND):
    out = []
    for s in seeds:
        st = {k: np.asfortranarray(v.copy()) for k, v in sp.state.items()}
        g = {k: np.asfortranarray(v) for k, v in sp.geom.items()}
        info = oracle.spicule_run(st, g, np.asfortranarray(sp.x), sp.margin, sp.gm, tend, nstop, 0.0, 0, s)
        out.append((st, info))
    return out


@pytest.mark.parametrize("bt", [64, 256, 1024])
def test_loop_realizations(oracle, dev, bt):
    print(f"spicule loop bt={bt}", flush=True)
    sp = spicule.setup()
    seeds = [1, 20160809, 123456789]
    nstop = 300
    ref = oracle_runs(oracle, sp, seeds, nstop)
    runs = spicule.SpiculeRuns(sp, seeds, device=dev)
    runs.run(nstop=nstop, block_threads=bt)
    torch.cuda.synchronize()
    for r, (st, (t, ns, merr, seed)) in enumerate(ref):
        info = runs.info(r)
        assert (info["ns"], info["merr"], info["seed"]) == (ns, merr, seed)
        assert info["t"] == t
        got = runs.state(r)
        for k in F:
            assert_bitwise(got[k], st[k], f"run {r} {k}")


def test_loop_resumes_bitwise(oracle, dev):
    sp = spicule.setup()
    runs = spicule.SpiculeRuns(sp, [4242, 77], device=dev)
    runs.run(nstop=120)
    runs.run(nstop=250)
    one = spicule.SpiculeRuns(sp, [4242, 77], device=dev)
    one.run(nstop=250)
    torch.cuda.synchronize()
    for r in range(2):
        assert runs.info(r) == one.info(r)
        a, b = runs.state(r), one.state(r)
        for k in F:
            assert_bitwise(a[k], b[k], k)


def test_loop_snapshots(oracle, dev):
    """Snapshots are taken where main.f writes its DAC records: after the
    steps at which int(t/dtout) changes."""
    sp = spicule.setup()
    dtout = 5.0e-5        # dt is ~2e-6 at the start: a record every ~25 steps
    runs = spicule.SpiculeRuns(sp, [31337], device=dev, nsnap_max=64, dtout=dtout)
    runs.run(nstop=200)
    torch.cuda.synchronize()
    times, states = runs.snapshots(0)
    st = {k: np.asfortranarray(v.copy()) for k, v in sp.state.items()}
    g = {k: np.asfortranarray(v) for k, v in sp.geom.items()}
    x = np.asfortranarray(sp.x)
    t, ns, seed = 0.0, 0, 31337
    want = []
    for _ in range(200):
        tp = t
        t, ns, merr, seed = oracle.spicule_run(st, g, x, sp.margin, sp.gm, spicule.TEND, ns + 1, t, ns, seed)
        if int(tp / dtout) < int(t / dtout):
            want.append((t, {k: st[k].copy() for k in F}))
    assert len(want) == len(times) > 3
    for (tw, sw), tg, sg in zip(want, times, states):
        assert tg == tw
        for q, k in enumerate(F):
            assert_bitwise(sg[q], sw[k], k)


def test_loop_per_run_geometry(oracle, dev):
    """Realizations with their own set-up (model.f's onebeta: field strength
    and flux-tube cross-section)."""
    sps = [spicule.setup(onebeta=b) for b in (1.0, 0.25)]
    runs = spicule.SpiculeRuns(sps, [5, 6], device=dev)
    runs.run(nstop=150)
    torch.cuda.synchronize()
    for r, (sp, s) in enumerate(zip(sps, [5, 6])):
        (st, info), = oracle_runs(oracle, sp, [s], 150)
        got = runs.state(r)
        assert runs.info(r)["t"] == info[0]
        for k in F:
            assert_bitwise(got[k], st[k], f"run {r} {k}")


def test_loop_dtmin_stop(dev):
    """cfl_m's merr = 9001 when dt < dtmin: the loop stops at ns = 1 and the
    state is not advanced."""
    sp = spicule.setup()
    runs = spicule.SpiculeRuns(sp, [9], device=dev, dtmin=1.0)
    runs.run(nstop=100)
    torch.cuda.synchronize()
    info = runs.info(0)
    assert info["merr"] == 9001 and info["ns"] == 1 and info["t"] == 0.0 and info["seed"] == 9
    got = runs.state(0)
    for k in F:
        assert_bitwise(got[k], sp.state[k], k)


def test_ensemble_long_run_properties(dev):
    """2000 steps of 64 realizations: equal seeds give equal runs, different
    seeds differ, everything stays finite and above the floor."""
    sp = spicule.setup()
    seeds = [1000 + (k % 32) for k in range(64)]
    runs = spicule.SpiculeRuns(sp, seeds, device=dev)
    runs.run(nstop=2000)
    torch.cuda.synchronize()
    a, b, c = runs.state(0), runs.state(32), runs.state(1)
    for k in F:
        assert_bitwise(a[k], b[k], k)
        assert np.isfinite(a[k]).all()
    assert not np.array_equal(a["vy"], c["vy"])
    assert (a["ro"] >= spicule.FLOOR).all() and (a["pr"] >= spicule.FLOOR).all()
    assert runs.info(0)["ns"] == 2000
