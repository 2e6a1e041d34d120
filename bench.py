#!/usr/bin/env python3
"""Benchmark of the CANS MHD time step on one MI355X.

Workload (default): 2D Orszag-Tang vortex, 1024^2 interior cells, periodic
box [0,2pi)^2, gamma = 5/3, CFL safety 0.4, artificial viscosity 3.0, FP64
modified Lax-Wendroff (the cans2d mlw_m path).  One step is one pass of the
CANS main loop over the whole grid: cfl_m -> t = t + dt -> mlw_m -> bnd
(periodic ghost fill and the next step's cfl_m are fused into the engine
kernel; dt, t and the stop test stay on the device).

Printed: one JSON line with zone-updates/s (interior cells x steps / s), the
roofline of the engine kernel (algorithmic bytes per launch / measured
average launch duration, HIP events on the launch stream) and a CPU baseline:
the reference Fortran routines compiled from source (oracle/_ref) run on the
host for a bounded number of steps of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "zone-updates/s on 1 MI355X (FP64 plain vs FP32-mixed/batched), %HBM peak, adj/fwd"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 vector (FMA counted as two operations)
# compulsory HBM traffic of one 2D zone-update: 7 FP64 fields read + 7 written
B_ALG_2D = 112


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2000)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--n", type=int, default=1024, help="interior cells per side")
    p.add_argument("--cpu-steps", type=int, default=100, help="reference steps timed on the host (one core)")
    p.add_argument("--cpu-procs", type=int, default=0,
                   help="independent reference processes of the all-cores figure (0: the host's usable cores, at most 16)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-extra", action="store_true", help="only the main workload (no 4096^2 line)")
    p.add_argument("--c2-steps", type=int, default=10, help="timed steps of the 3D blast figures (0: skip)")
    p.add_argument("--c2-l1-steps", type=int, default=500,
                   help="steps after which the FP32-flux vs FP64 L1 per field is taken")
    p.add_argument("--c2-n", type=int, default=384, help="interior cells per side of the 3D blast")
    p.add_argument("--tubes", type=int, default=4096, help="batched 1D tubes to t=0.1 (0: skip)")
    p.add_argument("--runs2d", type=int, default=1024, help="batched uneven 2D runs to t=0.2 (0: skip)")
    p.add_argument("--no-adjoint", dest="adjoint", action="store_false", help="skip the adjoint timing")
    p.add_argument("--spicule-runs", type=int, default=1024,
                   help="realizations of the UTRIP spicule ensemble figure (0: skip the spicule figures)")
    p.add_argument("--spicule-steps", type=int, default=100000, help="steps of the single UTRIP spicule run")
    p.add_argument("--copy-mib", type=int, default=0,
                   help="only run 5 device-to-device copies of this many MiB (the byte-counter calibration "
                        "for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE) and exit")
    return p.parse_args()


def copy_calibration(mib: int):
    """A plain copy kernel of known size: the HBM counters of its dispatches
    scale the counters read for the engine kernel."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.arange(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(5):
        torch.mul(src, 1.0, out=dst)   # an elementwise kernel: 8 B read + 8 B written per element
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "dispatches": 5}), flush=True)


def profile_value(name: str, key: str):
    """A KEY=value line of a committed profile summary (None if absent)."""
    f = ROOT / "profiles" / name
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith(key + "="):
            try:
                return float(line.split("=", 1)[1])
            except ValueError:
                return None
    return None


def traffic_from_profile(name: str):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary."""
    f = ROOT / "profiles" / name
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("TRAFFIC_BYTES_PER_LAUNCH"):
            try:
                return float(line.split("=")[1])
            except (IndexError, ValueError):
                return None
    return None


def host_info():
    """Core count and model of the host the CPU legs run on."""
    model = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"nproc": os.cpu_count(), "usable_cores": usable, "cpu_model": model}


def _ot_ref_setup(n: int):
    from oracle.oracle import Ref
    from utrip_amd import problems
    pb = problems.orszag_tang_2d(n=n, margin=1)
    return Ref("2d"), {k: v.copy(order="F") for k, v in pb.fields.items()}, pb.grid, pb.params


def _ot_ref_step(R, f, g, pp):
    """cfl_m + mlw_m + bdperx/bdpery of the reference build (cans2d main.f order)."""
    import ctypes as C
    ix, jx = f["ro"].shape
    names = ("ro", "pr", "vx", "vy", "bx", "by", "az")
    dt = C.c_double()
    merr = C.c_int()
    R.call("cfl_m", dt, pp["safety"], pp["dtmin"], merr, pp["gm"], *[f[k] for k in names[:6]], g["dx"], ix,
           g["dy"], jx)
    R.call("mlw_m", *[f[k] for k in names], dt.value, pp["qav"], pp["gm"], g["dx"], g["dxm"], ix, g["dy"],
           g["dym"], jx)
    for k in names:
        R.call("bdperx", 1, 1, f[k], ix, jx)
    for k in names:
        R.call("bdpery", 1, 1, f[k], ix, jx)


def _ot_ref_worker(n: int, nsteps: int):
    """One process of the all-cores figure: its own copy of the n^2 problem,
    1 warm-up step, then nsteps timed steps; returns the step times."""
    R, f, g, pp = _ot_ref_setup(n)
    _ot_ref_step(R, f, g, pp)
    out = []
    for _ in range(nsteps):
        t0 = time.perf_counter()
        _ot_ref_step(R, f, g, pp)
        out.append(time.perf_counter() - t0)
    return out


def cpu_baseline(n: int, nsteps: int, procs: int):
    """The reference Fortran (oracle/_ref/libcans2d_ref.so, built from the
    reference sources with amdflang -O2 -ffp-contract=off) on the host:
      one core -- cfl_m + mlw_m + bdperx/bdpery per step, 1 warm-up step,
                  median of nsteps step times;
      all cores -- `procs` independent processes, each running its own copy
                  of the same problem (the reference has no threads; this is
                  the host's whole-chip throughput on independent runs),
                  10 timed steps each, the sum of their rates."""
    from oracle.oracle import REF_LIBS
    if not REF_LIBS["2d"].exists():
        return None
    times = _ot_ref_worker(n, nsteps)
    med = statistics.median(times)
    zones = n * n
    info = host_info()
    out = {"value": zones / med, "unit": "zone-updates/s", "cores": 1, "kind": "reference",
           "sample": f"{nsteps} timed steps (after 1 warm-up) of the same {n}^2 Orszag-Tang MLW loop, "
                     f"median step {med:.3f} s; reference cfl_m/mlw_m/bdperx/bdpery compiled from source "
                     f"with amdflang -O2 -ffp-contract=off, single thread", **info}
    procs = procs or min(16, info["usable_cores"] or 1)
    if procs > 1:
        import multiprocessing as mp
        ctx = mp.get_context("spawn")
        t0 = time.perf_counter()
        with ctx.Pool(procs) as pool:
            res = pool.starmap(_ot_ref_worker, [(n, 10)] * procs)
        wall = time.perf_counter() - t0
        rate = sum(zones / statistics.median(r) for r in res)
        out["all_cores"] = {"value": rate, "unit": "zone-updates/s", "processes": procs,
                            "median_step_s": statistics.median([statistics.median(r) for r in res]),
                            "wall_s": wall,
                            "sample": f"{procs} independent processes, each its own {n}^2 problem, 10 timed steps "
                                      f"after 1 warm-up; sum of the per-process rates"}
    return out


def blast3d_figures(n: int, steps: int, l1_steps: int = 500, early: int = 100):
    """The 384^3 periodic MHD blast wave (rho=1, p=0.1, p=10 inside r<0.1,
    B_HL=(1,1,0)/sqrt(2)) on the resident run: 3D MLW (FP64), Roe-MUSCL and
    HLLD in FP64 and with FP32 face fluxes.  Each engine is timed over
    `steps` steps after 2 warm-up steps; the Roe and HLLD runs then continue
    to l1_steps steps (the workload's run length), where the relative L1
    difference per field between the FP32-flux and FP64 states is taken
    (density also after `early` steps).  A control run of each, FP64 from an
    initial density perturbed by 1e-7 relative (seeded), measures how far
    the run itself amplifies a perturbation of that size by then."""
    import gc

    import numpy as np
    import torch

    from utrip_amd import cans, problems
    dev = torch.device("cuda:0")
    F3 = cans.FIELDS[3]
    out = {"cells": n ** 3, "steps": steps, "warmup": 2, "l1_after_steps": l1_steps}
    with_l1 = l1_steps > max(early, steps + 2)
    variants = (("mlw", "fp64"), ("roe", "fp64"), ("roe", "fp32flux"), ("roe", "fp64_perturbed"),
                ("hlld", "fp64"), ("hlld", "fp32flux"), ("hlld", "fp64_perturbed"))
    if not with_l1:
        variants = tuple(v for v in variants if v[1] != "fp64_perturbed")
    finals, early_rho = {}, {}
    for engine, var in variants:
        margin = 2 if engine == "mlw" else 3
        pb = problems.blast_3d(n=n, margin=margin, gaussian_b=(engine != "hlld"))
        st = [torch.from_numpy(np.ascontiguousarray(pb.fields[k].T)).to(dev) for k in F3]
        gd = {k: torch.from_numpy(pb.grid[k]).to(dev) for k in ("dx", "dxm", "dy", "dym", "dz", "dzm")}
        del pb
        if var == "fp64_perturbed":
            gen = torch.Generator(device="cpu").manual_seed(3)
            xi = torch.rand(st[0].shape, generator=gen, dtype=torch.float64).to(dev) * 2 - 1
            st[0].mul_(1 + 1e-7 * xi)
            del xi
        eng = {"mlw": cans.ENGINE_MLW, "roe": cans.ENGINE_ROE, "hlld": cans.ENGINE_HLLD}[engine]
        run = cans.Run(3, eng, st, gd, 5.0 / 3.0, 0.4, qav=3.0, mu=1.0, margin=margin, bc=cans.BD_PER,
                       mixed=(var == "fp32flux"))
        run.reset(dtmin=1e-10)
        run.steps(2)
        ev = cans.EventSet(2)
        torch.cuda.synchronize()
        ev.record(0)
        run.steps(steps)
        ev.record(1)
        torch.cuda.synchronize()
        ms = ev.elapsed_ms(0, 1)
        name = engine + "_" + var
        if var != "fp64_perturbed":
            out[name] = {"zone_updates_per_s": n ** 3 * steps / (ms / 1e3), "ms_per_step": ms / steps,
                         "hbm_frac_alg": 128 * n ** 3 * steps / (ms / 1e3) / 1e9 / HBM_PEAK_GBS}
        total = steps + 2
        if engine != "mlw" and with_l1:
            sl = (slice(margin, -margin),) * 3
            run.steps(early - total)
            cur, _ = run.state()
            early_rho[name] = cur[0][sl].clone()
            run.steps(l1_steps - early)
            total = l1_steps
        cur, info = run.state()
        assert info["merr"] == 0 and info["nexec"] == total, info
        print(f"bench: blast {name} done", file=sys.stderr, flush=True)
        if engine != "mlw" and with_l1:
            out.setdefault(name, {})["t_at_l1"] = info["t"]
            finals[name] = [t.clone() for t in cur]   # kept on the GPU (3.7 GB per run at 384^3)
        del run, st, gd, cur
        gc.collect()
        torch.cuda.empty_cache()

    def rel(b, a):
        return {k: float((x - y).abs().sum() / y.abs().sum()) if float(y.abs().sum()) > 0
                else float((x - y).abs().sum()) for k, x, y in zip(F3, b, a)}

    for engine in ("roe", "hlld"):
        out[engine + "_fp32flux_speedup"] = out[engine + "_fp64"]["ms_per_step"] / out[engine + "_fp32flux"]["ms_per_step"]
        for var in ("fp64", "fp32flux"):
            ops = profile_value("engines3d_blast384_pmc.txt", f"FP64_OPS_PER_ZONE_UPDATE_{engine.upper()}_{var.upper()}")
            ops32 = profile_value("engines3d_blast384_pmc.txt", f"FP32_OPS_PER_ZONE_UPDATE_{engine.upper()}_{var.upper()}")
            zu = out[engine + "_" + var]["zone_updates_per_s"]
            if ops is not None:
                # FP64 arithmetic rate of the whole step (the op count per
                # zone-update from the committed counter profile, the rate
                # from this run) against the 78.6 TFLOP/s FP64 vector peak
                out[engine + "_" + var]["fp64_tflops"] = ops * zu / 1e12
                out[engine + "_" + var]["fp64_valu_frac"] = ops * zu / 1e12 / FP64_PEAK_TFLOPS
            if ops32 is not None:
                out[engine + "_" + var]["fp32_tflops"] = ops32 * zu / 1e12
        if not with_l1:
            continue
        a = finals[engine + "_fp64"]
        out[engine + "_rel_l1_fp32flux_vs_fp64"] = rel(finals[engine + "_fp32flux"], a)
        out[engine + "_rel_l1_fp64_perturbed_vs_fp64"] = rel(finals[engine + "_fp64_perturbed"], a)
        e0 = early_rho[engine + "_fp64"]
        out[engine + f"_rho_rel_l1_after_{early}_steps"] = {
            v: float((early_rho[engine + "_" + v] - e0).abs().sum() / e0.abs().sum())
            for v in ("fp32flux", "fp64_perturbed")}
    finals.clear()
    early_rho.clear()
    gc.collect()
    torch.cuda.empty_cache()
    return out


def c0_figures(no_cpu: bool = False, reps: int = 5):
    """The c0 workload: one Brio-Wu shock tube (nx = 1024, margin 2, free
    boundaries, gamma 2, safety 0.4, qav 3) run to t = 0.1 on the device
    clock (cfl_m, mlw_m, bdfrex launches per step, no host round trip);
    median of `reps` whole runs after a warm-up run.  The CPU leg runs the
    same loop over the reference routines built from source (oracle/_ref,
    libcans1d_ref.so) from C (oracle's orref_loop_1d), one core, median of
    `reps` whole runs."""
    import numpy as np
    import torch

    from utrip_amd import cans, problems
    pb = problems.shock_tube_1d(nx=1024, margin=2)
    g, pp = pb.grid, pb.params

    def td(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")

    gd = {k: td(g[k]) for k in ("dx", "dxm")}
    ms = []
    issue = 2000                            # warm-up run: the clock stops it at t >= 0.1
    for rep in range(reps + 1):
        st = [td(pb.fields[k]) for k in cans.FIELDS[1]]
        run = cans.Run(1, cans.ENGINE_MLW, st, gd, pp["gm"], pp["safety"], qav=pp["qav"], margin=2, bc=0,
                       bx1d=td(pb.fields["bx"]), bxm1d=td(pb.fields["bxm"]))
        run.reset(tend=pp["tend"], dtmin=pp["dtmin"])
        ev = cans.EventSet(2)
        torch.cuda.synchronize()
        ev.record(0)
        run.steps(issue)
        ev.record(1)
        torch.cuda.synchronize()
        _, info = run.state()
        assert info["merr"] == 0 and info["t"] >= pp["tend"], info
        if rep:
            assert info["nexec"] == issue, info
            ms.append(ev.elapsed_ms(0, 1))
        issue = info["nexec"]               # timed runs issue exactly the steps the run takes
    nsteps = info["nexec"]
    med = statistics.median(ms)
    out = {"cells": 1024, "steps": nsteps, "t": info["t"],
           "launch_per_step": {"ms": med, "us_per_step": med * 1e3 / nsteps,
                               "zone_updates_per_s": 1024 * nsteps / (med / 1e3),
                               "note": "cans_run: cfl_m, clock, mlw_m and bdfrex launches per step"}}
    # the same run as one persistent workgroup (the batched-tube kernel with
    # one tube): the whole loop on the device, one barrier per step
    tube = {k: pb.fields[k] for k in ("ro", "pr", "vx", "vy", "by", "bx", "bxm")}
    tube["dx"], tube["dxm"] = g["dx"], g["dxm"]
    b = cans.Batch1D([tube], pp["gm"], pp["qav"], pp["safety"], 2)
    init = [t.clone() for t in b.f]
    b.tune(probe_steps=50)
    ms = []
    for rep in range(reps + 1):
        for t_, i0 in zip(b.f, init):
            t_.copy_(i0)
        ev = cans.EventSet(2)
        torch.cuda.synchronize()
        ev.record(0)
        b.run(tend=pp["tend"])
        ev.record(1)
        torch.cuda.synchronize()
        if rep:
            ms.append(ev.elapsed_ms(0, 1))
    assert int(b.ns[0]) == nsteps and int(b.merr[0]) == 0 and float(b.t[0]) == info["t"]
    med = statistics.median(ms)
    out["persistent"] = {"ms": med, "us_per_step": med * 1e3 / nsteps, "zone_updates_per_s": 1024 * nsteps / (med / 1e3),
                         "block_threads": b.shape[0]}
    out["zone_updates_per_s"] = max(out["persistent"]["zone_updates_per_s"],
                                    out["launch_per_step"]["zone_updates_per_s"])
    if not no_cpu:
        from oracle.oracle import REF_LIBS, Oracle, Ref
        if REF_LIBS["1d"].exists():
            o, R = Oracle(), Ref("1d")
            secs = []
            for rep in range(reps + 1):
                f = {k: v.copy() for k, v in pb.fields.items()}
                t0 = time.perf_counter()
                t, ns, merr = o.ref_loop_1d(R, f, f["bx"], f["bxm"], g["dx"], g["dxm"], 2, pp["gm"], pp["safety"],
                                            pp["qav"], pp["dtmin"], pp["tend"], 10 ** 6)
                if rep:
                    secs.append(time.perf_counter() - t0)
            assert merr == 0 and ns == nsteps, (ns, nsteps)
            sec = statistics.median(secs)
            out["cpu_baseline"] = {"value": 1024 * ns / sec, "unit": "zone-updates/s", "cores": 1,
                                   "kind": "reference", "s": sec,
                                   "sample": f"the whole run ({ns} steps to t=0.1), median of {reps} runs; "
                                             "reference cfl_m/mlw_m/bdfrex (amdflang -O2 -ffp-contract=off) "
                                             "called from a C loop"}
            out["gpu_vs_cpu"] = out["zone_updates_per_s"] / out["cpu_baseline"]["value"]
    return out


def tubes_figures(ntube: int = 4096, tend: float = 0.1):
    """The batched-tube workload: ntube seeded Brio-Wu tubes (nx 256..4096,
    beta 1e-2..1e2, tests/golden/tubes_4096.npz), each run to t >= tend
    with its own cfl dt in one persistent kernel; FP64 (bit-exact path) and
    FP32 fluxes, after the launch-shape autotuner.  zone-updates = sum over
    tubes of nx x steps."""
    import numpy as np
    import torch

    from utrip_amd import cans, problems
    nx, beta, s = problems.tube_set(ntube)
    tubes = problems.tubes(nx, s, margin=2)
    out = {"tubes": ntube, "cells": int(nx.sum()), "tend": tend, "seed": problems.C3_SEED}
    finals = {}
    for mixed in (False, True):
        b = cans.Batch1D(tubes, 2.0, 3.0, 0.4, 2, mixed=mixed)
        init = [t.clone() for t in b.f]
        bt, bpc, tune_ms = b.tune(probe_steps=20)
        g = b.graph(tend=tend, init=init)
        g.launch()                      # warm-up replay
        torch.cuda.synchronize()
        ev = cans.EventSet(2)
        ev.record(0)
        g.launch()                      # timed replay: the whole batch from the initial state
        ev.record(1)
        torch.cuda.synchronize()
        ms = ev.elapsed_ms(0, 1)
        ns = b.ns.cpu().numpy().astype(np.int64)
        merr = b.merr.cpu().numpy()
        assert (merr == 0).all() and (b.t.cpu().numpy() >= tend).all()
        zu = int((nx * ns).sum())
        name = "fp32flux" if mixed else "fp64"
        out[name] = {"zone_updates": zu, "ms": ms, "zone_updates_per_s": zu / (ms / 1e3),
                     "steps_min": int(ns.min()), "steps_max": int(ns.max()), "launch_shape": [bt, bpc],
                     "hbm_frac_alg": 112 * zu / (ms / 1e3) / 1e9 / HBM_PEAK_GBS}
        finals[name] = ([t.clone() for t in b.f], b)
        if not mixed:
            ns_fp64 = ns
            ns_fp64_dev = b.ns.clone()
    a, _ = finals["fp64"]
    c, bb = finals["fp32flux"]

    def rel_rho(f):
        r = []
        for k in range(ntube):
            sl = slice(int(bb.off_host[k]), int(bb.off_host[k] + bb.ix_host[k]))
            r.append(float((f[0][sl] - a[0][sl]).abs().sum() / a[0][sl].abs().sum()))
        return np.array(r)

    rel = rel_rho(c)
    ns32 = bb.ns.cpu().numpy()
    same = ns32 == ns_fp64
    # the FP32-flux run again from the initial state, each tube for exactly
    # the steps its FP64 run took
    for t_, i0 in zip(bb.f, init):
        t_.copy_(i0)
    bb.run_steps(ns_fp64_dev)
    torch.cuda.synchronize()
    rel_same = rel_rho(bb.f)
    out["rho_rel_l1_fp32flux_vs_fp64"] = {
        "same_steps": {"max": float(rel_same.max()), "median": float(np.median(rel_same))},
        "to_tend": {"max": float(rel.max()), "median": float(np.median(rel)),
                    "tubes_other_step_count": int((~same).sum()),
                    "max_over_tubes_same_step_count": float(rel[same].max())}}
    # the alternative the batch replaces: the plain single-run engine looped
    # over the tubes (one resident 1D run per tube, cfl + mlw_m + bdfre per
    # step), timed on the first `nloop` tubes of the set for the same steps
    nloop = 16
    zu = 0
    ms = 0.0
    for k in range(nloop):
        d = tubes[k]
        st = [torch.from_numpy(np.ascontiguousarray(d[q])).to("cuda:0") for q in cans.FIELDS[1]]
        gd = {"dx": torch.from_numpy(d["dx"]).to("cuda:0"), "dxm": torch.from_numpy(d["dxm"]).to("cuda:0")}
        run = cans.Run(1, cans.ENGINE_MLW, st, gd, 2.0, 0.4, qav=3.0, margin=2, bc=0,
                       bx1d=torch.from_numpy(d["bx"]).to("cuda:0"), bxm1d=torch.from_numpy(d["bxm"]).to("cuda:0"))
        run.reset(tend=tend)
        ev = cans.EventSet(2)
        torch.cuda.synchronize()
        ev.record(0)
        run.steps(int(ns_fp64[k]))
        ev.record(1)
        torch.cuda.synchronize()
        _, info = run.state()
        assert info["merr"] == 0 and info["nexec"] == int(ns_fp64[k]), info
        zu += int(nx[k]) * int(ns_fp64[k])
        ms += ev.elapsed_ms(0, 1)
    out["looped_single_runs"] = {"tubes": nloop, "zone_updates": zu, "ms": ms,
                                 "zone_updates_per_s": zu / (ms / 1e3),
                                 "batched_fp64_over_looped": out["fp64"]["zone_updates_per_s"] / (zu / (ms / 1e3))}
    return out


def batch2d_figures(nrun: int = 1024, tend: float = 0.2):
    """The batched 2D workload: nrun periodic Orszag-Tang runs of uneven size
    (nx, ny uniform in 32..256, pressure scale 0.1..10, seeded:
    problems.ot_run_set), each running the cans2d main loop to its own
    t >= tend in one persistent kernel (cans2d_batch_run), FP64, against the
    same runs looped one after another through the single-run engine.  The
    launch shape comes from the autotuner (cans2d_batch_tune); the timed run
    is a replay of the batch's hipGraph."""
    import numpy as np
    import torch

    from utrip_amd import cans, problems
    nx, ny, ps = problems.ot_run_set(nrun)
    runs = []
    for k in range(nrun):
        pb = problems.orszag_tang_box(int(nx[k]), int(ny[k]), 1, float(ps[k]))
        r = dict(pb.fields)
        r.update({q: pb.grid[q] for q in ("dx", "dxm", "dy", "dym")})
        runs.append(r)
    pp = pb.params
    cells = (nx * ny).astype(np.int64)
    out = {"runs": nrun, "cells": int(cells.sum()), "tend": tend, "n_range": [32, 256]}
    b = cans.Batch2D(runs, pp["gm"], pp["qav"], pp["safety"], 1)
    init = [t.clone() for t in b.f]
    bt, bpc, tune_ms = b.tune(probe_steps=10)
    g = b.graph(tend=tend, init=init)
    g.launch()                      # warm-up replay
    torch.cuda.synchronize()
    ev = cans.EventSet(2)
    ev.record(0)
    g.launch()                      # timed replay: the whole batch from the initial state
    ev.record(1)
    torch.cuda.synchronize()
    ms = ev.elapsed_ms(0, 1)
    ns = b.ns.cpu().numpy().astype(np.int64)
    assert (b.merr.cpu().numpy() == 0).all() and (b.t.cpu().numpy() >= tend).all()
    zu = int((cells * ns).sum())
    out["fp64"] = {"zone_updates": zu, "ms": ms, "zone_updates_per_s": zu / (ms / 1e3), "launch_shape": [bt, bpc],
                   "tune_probe_ms": tune_ms, "steps_min": int(ns.min()), "steps_max": int(ns.max()),
                   "hbm_frac_alg": B_ALG_2D * zu / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                   "note": "autotuned launch shape; the hipGraph of the whole batch run, replayed"}
    # the alternative the batch replaces: each run through the resident
    # single-run loop (MlwRun2D, one fused launch per step), one after
    # another, for the steps the batch took; timed on the first `nloop` runs
    nloop = 16
    zl, msl = 0, 0.0
    for k in range(nloop):
        st = [torch.from_numpy(np.ascontiguousarray(runs[k][q].T)).to("cuda:0") for q in cans.FIELDS[2]]
        gd = [torch.from_numpy(runs[k][q]).to("cuda:0") for q in ("dx", "dxm", "dy", "dym")]
        run = cans.MlwRun2D(st, *gd, pp["qav"], pp["gm"], pp["safety"])
        run.reset(tend=tend, dtmin=pp["dtmin"])
        ev = cans.EventSet(2)
        torch.cuda.synchronize()
        ev.record(0)
        run.steps(int(ns[k]))
        ev.record(1)
        torch.cuda.synchronize()
        msl += ev.elapsed_ms(0, 1)
        _, info = run.state()
        assert info["nexec"] == int(ns[k]) and info["t"] == float(b.t[k])
        zl += int(cells[k] * ns[k])
        del run, st, gd
    out["looped_single_runs"] = {"runs": nloop, "zone_updates": zl, "ms": msl, "zone_updates_per_s": zl / (msl / 1e3),
                                 "batched_over_looped": out["fp64"]["zone_updates_per_s"] / (zl / (msl / 1e3))}
    return out


def m3_figures(n: int = 1024, steps: int = 100):
    """The three-component engine (cans2d/hdmlw/mlw_m3.f with cfl_m3 and
    bdperx/bdpery, CANS_ENGINE_MLW3 of the resident loop) on the 2.5D
    Orszag-Tang vortex at n^2: the engine kernel's average launch (HIP events
    around it every step) against 144 B per zone-update (9 fields read and
    written), and the whole step (cfl pass + engine + boundary passes)."""
    import numpy as np
    import torch

    from utrip_amd import cans, problems
    pb = problems.orszag_tang_25d(n=n, margin=1)
    names = cans.FIELDS_M3[2]

    def td(a):
        return torch.from_numpy(np.ascontiguousarray(a.T)).to("cuda:0")

    st = [td(pb.fields[k]) for k in names]
    g = {k: torch.from_numpy(pb.grid[k]).to("cuda:0") for k in ("dx", "dxm", "dy", "dym")}
    pp = pb.params
    run = cans.Run(2, cans.ENGINE_MLW3, st, g, pp["gm"], pp["safety"], qav=pp["qav"], margin=1, bc=7)
    run.reset(dtmin=pp["dtmin"])
    run.steps(5)
    ev = cans.EventSet(2 * steps + 2)
    torch.cuda.synchronize()
    ev.record(2 * steps)
    run.steps(steps, events=ev)
    ev.record(2 * steps + 1)
    torch.cuda.synchronize()
    total = ev.elapsed_ms(2 * steps, 2 * steps + 1)
    kern = statistics.fmean(ev.elapsed_ms(2 * k, 2 * k + 1) for k in range(steps))
    _, info = run.state()
    assert info["merr"] == 0 and info["nexec"] == steps + 5, info
    z = n * n
    return {"cells": z, "steps": steps, "ms_per_step": total / steps, "zone_updates_per_s": z * steps / (total / 1e3),
            "engine_avg_us": kern * 1e3, "engine_hbm_frac": 144 * z / (kern / 1e3) / 1e9 / HBM_PEAK_GBS}


def roe2d_figures(n: int = 1024, steps: int = 100):
    """The same Orszag-Tang vortex at n^2 with the 2D Roe-MUSCL engine
    (cans2d/hdroe/roe_m.f; margin 3 as the reference drivers use, periodic
    bdper, cfl_m each step): whole-step throughput and the engine's average
    time (HIP events around it every step) against 112 B per zone-update."""
    import numpy as np
    import torch

    from utrip_amd import cans, problems
    pb = problems.orszag_tang_2d(n=n, margin=3)

    def td(a):
        return torch.from_numpy(np.ascontiguousarray(a.T)).to("cuda:0")

    st = [td(pb.fields[k]) for k in cans.FIELDS[2]]
    g = {k: torch.from_numpy(pb.grid[k]).to("cuda:0") for k in ("dx", "dxm", "dy", "dym")}
    pp = pb.params
    run = cans.Run(2, cans.ENGINE_ROE, st, g, pp["gm"], pp["safety"], qav=pp["qav"], margin=3, bc=7)
    run.reset(dtmin=pp["dtmin"])
    run.steps(5)
    ev = cans.EventSet(2 * steps + 2)
    torch.cuda.synchronize()
    ev.record(2 * steps)
    run.steps(steps, events=ev)
    ev.record(2 * steps + 1)
    torch.cuda.synchronize()
    total = ev.elapsed_ms(2 * steps, 2 * steps + 1)
    kern = statistics.fmean(ev.elapsed_ms(2 * k, 2 * k + 1) for k in range(steps))
    _, info = run.state()
    assert info["merr"] == 0 and info["nexec"] == steps + 5, info
    z = n * n
    return {"cells": z, "steps": steps, "ms_per_step": total / steps, "zone_updates_per_s": z * steps / (total / 1e3),
            "engine_avg_us": kern * 1e3, "engine_hbm_frac": B_ALG_2D * z / (kern / 1e3) / 1e9 / HBM_PEAK_GBS}


def spicule_cpu(steps: int):
    """CPU leg of the spicule figures: the reference's own UTRIP program
    (cans1d/md_spicule/hdmlw main.f with its model, bnd, bndsc, trqalfven and
    the reference routines, built from source by oracle/Makefile into
    oracle/_ref/spicule_ref), its whole 100000-step run in a scratch folder,
    one core.  Where that program is not built, the C restatement's loop
    (oracle/liboracle.so) for `steps` steps instead."""
    import subprocess
    import tempfile

    exe = ROOT / "oracle" / "_ref" / "spicule_ref"
    if exe.exists():
        with tempfile.TemporaryDirectory() as d:
            t0 = time.perf_counter()
            p = subprocess.run([str(exe)], cwd=d, capture_output=True, text=True, timeout=300)
            sec = time.perf_counter() - t0
        assert p.returncode == 0 and "normal stop" in p.stdout, p.stdout[-500:]
        ns = int(p.stdout.split("stop ")[1].split("step=")[1].split()[0])
        return {"steps": ns, "s": sec, "steps_per_s": ns / sec, "zone_updates_per_s": ns * 1024 / sec,
                "cores": 1, "kind": "reference",
                "sample": f"the reference md_spicule/hdmlw program built from source (amdflang -O2 "
                          f"-ffp-contract=off), its whole run: {ns} steps of 1024 cells, wall time incl. its DAC "
                          f"output"}
    import numpy as np

    from oracle.oracle import Oracle
    from utrip_amd import spicule
    sp = spicule.setup()
    st = {k: np.asfortranarray(v.copy()) for k, v in sp.state.items()}
    g = {k: np.asfortranarray(v) for k, v in sp.geom.items()}
    o = Oracle()
    t0 = time.perf_counter()
    o.spicule_run(st, g, np.asfortranarray(sp.x), sp.margin, sp.gm, spicule.TEND, steps, 0.0, 0, 1)
    sec = time.perf_counter() - t0
    return {"steps": steps, "s": sec, "steps_per_s": steps / sec, "zone_updates_per_s": steps * 1024 / sec,
            "cores": 1, "kind": "port", "sample": f"{steps} steps of the 1024-cell spicule loop from t=0"}


def spicule_figures(nrun: int, steps_single: int, cpu_steps: int = 10000, no_cpu: bool = False):
    """The UTRIP spicule model (md_spicule/hdmlw/main.f loop, 1024 cells,
    mlw_m_bg + Park-Miller torque + bnd + chkdav, FP64 bit-exact):
      single   -- one realization for the shipped run's length (100000 steps)
                  in one persistent workgroup; block size picked by a probe;
      ensemble -- nrun realizations (seeds 1..nrun), 2000 steps each, one
                  workgroup per realization.
    zone-updates = realizations x 1024 interior cells x steps; the HBM
    fraction counts 80 B per zone-update (5 fields read and written; the
    geometry is shared by all realizations and stays in cache)."""
    import torch

    from utrip_amd import cans, spicule
    sp = spicule.setup()
    out = {"cells": 1024}

    def timed(runs, nstop, bt):
        ev = cans.EventSet(2)
        torch.cuda.synchronize()
        ev.record(0)
        runs.run(nstop=nstop, block_threads=bt)
        ev.record(1)
        torch.cuda.synchronize()
        return ev.elapsed_ms(0, 1)

    probe = {}
    for bt in (256, 512, 1024):
        r = spicule.SpiculeRuns(sp, [1], device="cuda:0")
        timed(r, 200, bt)
        probe[bt] = timed(r, 2200, bt)
    bt = min(probe, key=probe.get)
    r = spicule.SpiculeRuns(sp, [1], device="cuda:0")
    ms = timed(r, steps_single, bt)
    info = r.info(0)
    assert info["merr"] == 0 and info["ns"] == steps_single, info
    out["single"] = {"steps": steps_single, "ms": ms, "us_per_step": ms * 1e3 / steps_single,
                     "t_reached": info["t"], "block_threads": bt,
                     "zone_updates_per_s": 1024 * steps_single / (ms / 1e3)}
    if nrun > 0:
        steps = 2000
        best = None
        for bte in (64, 128, 256):
            e = spicule.SpiculeRuns(sp, list(range(1, nrun + 1)), device="cuda:0")
            timed(e, 50, bte)
            m = timed(e, 250, bte)
            if best is None or m < best[1]:
                best = (bte, m)
        bte = best[0]
        e = spicule.SpiculeRuns(sp, list(range(1, nrun + 1)), device="cuda:0")
        ms = timed(e, steps, bte)
        assert int(e.merr.abs().max().item()) == 0 and int(e.ns.min().item()) == steps
        zu = nrun * 1024 * steps
        out["ensemble"] = {"realizations": nrun, "steps": steps, "ms": ms, "block_threads": bte,
                           "zone_updates_per_s": zu / (ms / 1e3),
                           "hbm_frac_alg": 80 * zu / (ms / 1e3) / 1e9 / HBM_PEAK_GBS}
    if not no_cpu and cpu_steps > 0:
        c = spicule_cpu(cpu_steps)
        out["cpu_baseline"] = c
        out["single"]["vs_cpu"] = out["single"]["zone_updates_per_s"] / c["zone_updates_per_s"]
        if "ensemble" in out:
            out["ensemble"]["vs_cpu"] = out["ensemble"]["zone_updates_per_s"] / c["zone_updates_per_s"]
    return out


def adjoint_figures(n: int = 512, nsteps: int = 200, ckpt: int = 20):
    """Gradient of L = sum (rho_final - rho_target)^2 dx dy over an n^2 OT run
    of nsteps MLW steps with respect to the initial velocities and a pressure
    scale: forward and hand-written reverse mode, wall time of each and the
    backward over the plain forward run (no recording).  Two storage modes:
    every state of the trajectory kept in device memory (the figure
    reported as adj_over_fwd), and a checkpoint every ckpt steps with the
    segments recomputed in the backward (under "checkpoints")."""
    out = None
    for mode in (0, ckpt):
        r = _adjoint_run(n, nsteps, mode)
        if out is None:
            out = r
            out["store"] = "every state kept (%.2f GB)" % ((nsteps + 1) * 7 * (n + 2) ** 2 * 8 / 1e9)
        else:
            out["checkpoints"] = {k: r[k] for k in ("checkpoint_every", "forward_ms", "backward_ms", "adj_over_fwd",
                                                    "dL_dbeta")}
    out["note"] = "adj_over_fwd = backward / plain forward (no recording) of the same 200 steps"
    return out


def _adjoint_run(n: int, nsteps: int, ckpt: int):
    import numpy as np
    import torch

    from utrip_amd import adjoint, cans, problems
    dev = torch.device("cuda:0")
    pb = problems.orszag_tang_2d(n=n, margin=1)
    names = ("ro", "pr", "vx", "vy", "bx", "by", "az")
    base = [torch.from_numpy(np.ascontiguousarray(pb.fields[k].T)).to(dev) for k in names]
    g = {k: torch.from_numpy(pb.grid[k]).to(dev) for k in ("dx", "dxm", "dy", "dym")}
    pp = pb.params
    solver = adjoint.PeriodicMlwAdjoint(g["dx"], g["dxm"], g["dy"], g["dym"], pp["qav"], pp["gm"], pp["safety"], 1,
                                        nsteps, ckpt=ckpt)
    gen = torch.Generator(device="cpu").manual_seed(7)
    xi = lambda: torch.randn((n, n), generator=gen, dtype=torch.float64).to(dev)
    vx0 = (base[2][1:-1, 1:-1] * (1 + 1e-3 * xi())).requires_grad_(True)
    vy0 = (base[3][1:-1, 1:-1] * (1 + 1e-3 * xi())).requires_grad_(True)
    beta = torch.ones((), dtype=torch.float64, device=dev, requires_grad=True)
    with torch.no_grad():
        target = adjoint.FinalDensity.apply(vx0 * (1 + 0.01 * xi()), vy0 * (1 + 0.01 * xi()), beta, solver, base)
    w = float(pb.grid["dx"][1] * pb.grid["dy"][1])
    # plain forward (no checkpoints) for reference
    run = cans.MlwRun2D([t.clone() for t in base], g["dx"], g["dxm"], g["dy"], g["dym"], pp["qav"], pp["gm"],
                        pp["safety"])
    run.reset()
    run.steps(nsteps)
    torch.cuda.synchronize()
    plains = []
    for _ in range(3):
        t0 = time.perf_counter()
        run.reset()
        run.steps(nsteps)
        torch.cuda.synchronize()
        plains.append(time.perf_counter() - t0)
    plain = statistics.median(plains)
    fwd, bwd = [], []
    for _ in range(3):
        for t in (vx0, vy0, beta):
            t.grad = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        L = ((adjoint.FinalDensity.apply(vx0, vy0, beta, solver, base) - target) ** 2).sum() * w
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        L.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        fwd.append(t1 - t0)
        bwd.append(t2 - t1)
    f, b = statistics.median(fwd), statistics.median(bwd)
    return {"cells": n * n, "steps": nsteps, "checkpoint_every": ckpt, "forward_plain_ms": plain * 1e3,
            "forward_ms": f * 1e3, "backward_ms": b * 1e3, "adj_over_fwd": b / plain,
            "adj_over_recording_fwd": b / f, "loss": float(L.detach()), "dL_dbeta": float(beta.grad)}


def adjoint_roe_figures(n: int = 512, nsteps: int = 50):
    """The reverse mode of the periodic 2D Roe-MUSCL loop (margin 3) on the
    n^2 Orszag-Tang vortex: gradient of L = sum (rho_final - rho_target)^2
    dx dy over nsteps steps with respect to the initial velocities and a
    pressure scale (every state kept); backward wall time against the plain
    forward run of the resident Roe engine for the same steps."""
    import numpy as np
    import torch

    from utrip_amd import adjoint, cans, problems
    dev = torch.device("cuda:0")
    m = 3
    pb = problems.orszag_tang_2d(n=n, margin=m)
    names = ("ro", "pr", "vx", "vy", "bx", "by", "az")
    base = [torch.from_numpy(np.ascontiguousarray(pb.fields[k].T)).to(dev) for k in names]
    gen = torch.Generator(device="cpu").manual_seed(7)
    xi = lambda shape: torch.randn(shape, generator=gen, dtype=torch.float64).to(dev)
    for q in (0, 1, 4, 5):
        base[q] = adjoint.periodic_fill(base[q] * (1 + 1e-3 * xi(base[q].shape)), m)
    dx = torch.from_numpy(pb.grid["dx"]).to(dev)
    dy = torch.from_numpy(pb.grid["dy"]).to(dev)
    pp = pb.params
    solver = adjoint.PeriodicRoeAdjoint(dx, dy, pp["gm"], pp["safety"], m, nsteps)
    vx0 = (base[2][m:-m, m:-m] * (1 + 1e-3 * xi((n, n)))).requires_grad_(True)
    vy0 = (base[3][m:-m, m:-m] * (1 + 1e-3 * xi((n, n)))).requires_grad_(True)
    beta = torch.ones((), dtype=torch.float64, device=dev, requires_grad=True)
    with torch.no_grad():
        target = adjoint.FinalDensity.apply(vx0 * (1 + 0.01 * xi((n, n))), vy0 * (1 + 0.01 * xi((n, n))), beta,
                                            solver, base)
    w = float(pb.grid["dx"][m] * pb.grid["dy"][m])
    grid = {"dx": dx, "dxm": dx, "dy": dy, "dym": dy}
    run = cans.Run(2, cans.ENGINE_ROE, [t.clone() for t in base], grid, pp["gm"], pp["safety"], margin=m,
                   bc=cans.BD_PER)
    plains = []
    for _ in range(3):
        run.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run.steps(nsteps)
        torch.cuda.synchronize()
        plains.append(time.perf_counter() - t0)
    plain = statistics.median(plains)
    fwd, bwd = [], []
    for _ in range(3):
        for t in (vx0, vy0, beta):
            t.grad = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        L = ((adjoint.FinalDensity.apply(vx0, vy0, beta, solver, base) - target) ** 2).sum() * w
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        L.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        fwd.append(t1 - t0)
        bwd.append(t2 - t1)
    f, b = statistics.median(fwd), statistics.median(bwd)
    return {"cells": n * n, "steps": nsteps, "forward_plain_ms": plain * 1e3, "forward_recording_ms": f * 1e3,
            "backward_ms": b * 1e3, "adj_over_fwd": b / plain, "loss": float(L.detach()),
            "dL_dbeta": float(beta.grad),
            "note": "every state and half step kept (cans2d_roe_record); adj_over_fwd = backward / plain forward"}


def progress(what: str) -> None:
    """A progress line on stderr (stdout carries only the JSON result)."""
    print(f"bench: {what} ({time.strftime('%H:%M:%S')})", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.copy_mib > 0:
        copy_calibration(args.copy_mib)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the library has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from utrip_amd import cans, problems

    dev = torch.device("cuda:0")

    def ot_run(n: int, K: int, warmup: int):
        """K timed steps of the OT n^2 periodic MLW loop after `warmup` steps,
        bracketed by two HIP events on the launch stream.  In this periodic
        path a step is exactly one kernel launch (mlw2d_kernel: the engine
        with the ghost fill and the next step's cfl_m fused in), so the timed
        region / K is the average launch duration of that kernel including
        the gaps between back-to-back launches (never shorter than the kernel,
        never longer than the step)."""
        pb = problems.orszag_tang_2d(n=n, margin=1)
        names = ("ro", "pr", "vx", "vy", "bx", "by", "az")
        g, pp = pb.grid, pb.params

        def td(a):
            return torch.from_numpy(np.ascontiguousarray(a.T)).to(dev)

        st = [td(pb.fields[k]) for k in names]
        gd = {k: td(v) for k, v in g.items()}
        run = cans.MlwRun2D(st, gd["dx"], gd["dxm"], gd["dy"], gd["dym"], pp["qav"], pp["gm"], pp["safety"])
        run.reset(dtmin=pp["dtmin"])
        mass0 = float(st[0][1:-1, 1:-1].sum())
        run.steps(warmup)
        ev = cans.EventSet(2)
        torch.cuda.synchronize()
        ev.record(0)
        run.steps(K)
        ev.record(1)
        torch.cuda.synchronize()
        total_ms = ev.elapsed_ms(0, 1)
        cur, info = run.state()
        assert info["merr"] == 0 and info["nexec"] == warmup + K, info
        mass1 = float(cur[0][1:-1, 1:-1].sum())
        return dict(pb=pb, total_ms=total_ms, avg_kern_s=total_ms / K / 1e3, info=info,
                    drift=(mass1 - mass0) / mass0)

    n = args.n
    K = args.steps
    r = ot_run(n, K, args.warmup)
    progress(f"OT {n}^2: {K} steps")
    pb, pp, total_ms, avg_kern_s, info = r["pb"], r["pb"].params, r["total_ms"], r["avg_kern_s"], r["info"]
    adj = adj_roe = None
    if args.adjoint:
        adj = adjoint_figures()
        progress("adjoint")
        adj_roe = adjoint_roe_figures()
        progress("Roe adjoint")
    m3 = roe2 = None
    if not args.no_extra:
        m3 = m3_figures()
        roe2 = roe2d_figures()
        progress("mlw_m3 and 2D Roe")
    spc = None
    if args.spicule_runs > 0 or args.spicule_steps > 0:
        spc = spicule_figures(args.spicule_runs, args.spicule_steps, no_cpu=args.no_cpu)
        progress("spicule")
    tb = None
    if args.tubes > 0:
        tb = tubes_figures(args.tubes)
        progress("tubes")
    b2 = None
    if args.runs2d > 0:
        b2 = batch2d_figures(args.runs2d)
        progress("batched 2D runs")
    c2 = None
    if args.c2_steps > 0:
        c2 = blast3d_figures(args.c2_n, args.c2_steps, args.c2_l1_steps)
        progress("3D blast")
    big = None
    if not args.no_extra:
        # the same loop at 4096^2: 0.94 GB of state per buffer set, beyond
        # the 256 MB Infinity Cache that holds the 1024^2 state
        rb = ot_run(4096, 100, 5)
        progress("OT 4096^2")
        zb = 4096 * 4096
        ach = B_ALG_2D * zb / rb["avg_kern_s"] / 1e9
        big = {"cells": zb, "steps": 100, "zone_updates_per_s": zb * 100 / (rb["total_ms"] / 1e3),
               "ms_per_step": rb["total_ms"] / 100, "avg_launch_us": rb["avg_kern_s"] * 1e6,
               "hbm_frac": ach / HBM_PEAK_GBS,
               "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": ach / HBM_PEAK_GBS, "kernel": "mlw2d_kernel",
                            "traffic": traffic_from_profile("mlw2d_kernel_ot4096_pmc.txt")}}
    zones = n * n
    value = zones * K / (total_ms / 1e3)
    achieved = B_ALG_2D * zones / avg_kern_s / 1e9
    traffic = traffic_from_profile("mlw2d_kernel_ot1024_pmc.txt")

    cpu = None
    if not args.no_cpu and args.cpu_steps > 0:
        cpu = cpu_baseline(n, args.cpu_steps, args.cpu_procs)
        progress("CPU baseline")
    c0 = None
    if not args.no_extra:
        c0 = c0_figures(no_cpu=args.no_cpu)
        progress("c0 shock tube")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": total_ms / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"2D Orszag-Tang vortex, {n}x{n} periodic, CANS mlw_m FP64 "
                               f"(cfl_m + mlw_m + bdper per step)",
                   "cells": zones, "gamma": pp["gm"], "safety": pp["safety"], "qav": pp["qav"],
                   "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "mlw2d_kernel", "bytes_per_zone_update": B_ALG_2D,
                     "avg_launch_us": avg_kern_s * 1e6},
        "cpu_baseline": cpu,
        "gpu_vs_cpu": (value / cpu["value"]) if cpu else None,
        "gpu_vs_cpu_all_cores": (value / cpu["all_cores"]["value"]) if cpu and "all_cores" in cpu else None,
        "c0_shock_tube": c0,
        "check": {"t": info["t"], "dt_last": info["dt"], "mass_rel_drift": r["drift"]},
        "ot4096_fp64": big,
        "blast384": c2,
        "tubes": tb,
        "runs2d": b2,
        "adjoint_ot512": adj,
        "adjoint_roe_ot512": adj_roe,
        "spicule": spc,
        "ot1024_mlw_m3": m3,
        "ot1024_roe": roe2,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
