#!/usr/bin/env python3
"""Benchmark of the MI355X hip_solver on its headline workload.

Workload (one GPU): the 8192 x 8192 FP64 dam break on [0,20]^2, flat bottom, reflecting
walls, h = 2 for x < 10 else 1 plus 1e-3 seeded uniform noise, u = v = 0, fixed dt from
CFL 0.45 at t = 0 (the reference's fixed-dt policy).  A "step" is one SSP-RK3 step of the
whole grid (three fused stage launches).  Inputs are resident in HBM before timing.

Prints ONE JSON line.  `value` = interior cell-updates per second / 1e9 over the timed
steps (wall clock around the replayed steps, synchronised).  `roofline` is priced from HIP
events recorded on the solver's stream around every stage launch in a second timed pass.
`cpu_baseline` times the reference's own OpenMP kernels (oracle/_ref) on the host on 10
steps of the same grid (one thread per available physical core, OMP_PROC_BIND=true,
OMP_PLACES=cores, in a child process); those steps are also compared bitwise with the GPU.
Its `c1` record is config 1 in full (256^2, g = 9.81, 1000 steps) on the host and on the
GPU, with the max relative L-inf difference of the final fields.

Also reported: `batched` (512-instance ensemble vs per-instance launches), `adjoint`
(1024^2, 2000 steps: adjoint wall time over forward wall time) and `serial_schemes` (the
reference SerialSolver's reconstruction x Riemann combinations at 4096^2, with the
reference's own modular code timed on a 1024^2 sample).

    python bench.py [--steps K] [--warmup W] [--n 8192] [--cpu-steps S] [--no-cpu] [--no-c1]
                    [--no-ensemble] [--no-adjoint] [--no-serial] [--serial-n N]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "shallow-water_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0            # MI355X HBM3E peak, GB/s
SHADER_CLOCK_GHZ = 2.4           # MI355X peak engine clock
# algorithmic HBM bytes per interior cell, stages 1..3: read U_s and B (and U^n for stages 2
# and 3), write U_{s+1}; with a bottom of +0.0 everywhere the stage kernels do not read B
BYTES_PER_CELL_STAGE = (56, 80, 80)
BYTES_PER_CELL_STAGE_FLAT_B = (48, 72, 72)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def mass(h, ng):
    return math.fsum(h[ng:-ng, ng:-ng].ravel().tolist())


def cpu_baseline(case, steps, dt, workload="dam_break"):
    """The reference's compiled OpenMP kernels (oracle/_ref; the oracle port when the
    reference harness is absent) on `steps` fixed-dt steps of the same grid, in a child
    process with one OpenMP thread per available physical core, OMP_PROC_BIND=true and
    OMP_PLACES=cores (scripts/leo3e/run_openmp.sh:111-114).  Returns the info dict and the
    sha256 of the final h, hu, hv (ghosts included) for the bitwise check."""
    from oracle import cpu_baseline as cb

    host = cb.host_cpus()
    r = cb.run_child(["--workload", workload, "--n", str(case["nx"]), "--steps", str(steps),
                      "--dt", float(dt).hex()], host["threads"], timeout=900)
    if r["rc"] != 0:
        raise RuntimeError("CPU baseline run failed")
    cells = case["nx"] * case["ny"]
    info = {
        "value": cells * steps / r["seconds"] / 1e9, "unit": "Gcell-updates/s",
        "cores": r["threads"], "kind": r["kind"],
        "sample": f"{case['nx']}x{case['ny']} {workload.replace('_', ' ')}, {steps} fixed-dt SSP-RK3 steps "
                  f"({'oracle/_ref: reference omp_kernels.hpp' if r['kind'] == 'reference' else 'oracle port'}"
                  f", OpenMP, g++ -O3 -ffp-contract=off), {r['seconds']:.2f} s",
        "threads": r["threads"], "binding": f"OMP_PROC_BIND={r['bind']} OMP_PLACES={r['places']}",
        "host": host,
    }
    return info, r["sha256"]


def c1_record(threads):
    """Config 1 as stated: 256^2 Gaussian pulse, g = 9.81, CFL 0.4 dt at t = 0, 1000
    fixed-dt steps.  The reference's kernels (the g = 9.81 harness build) on the host for
    the full run, and the GPU on the same dt; Gcell-updates/s of both and the max relative
    L-inf difference of the final fields (0 when bitwise equal)."""
    import tempfile

    import swe_hip
    from oracle import cpu_baseline as cb
    from swe_hip import workloads as W

    n, steps = 256, 1000
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "c1.npz")
        r = cb.run_child(["--workload", "gauss_pulse", "--n", str(n), "--steps", str(steps),
                          "--cfl", "0.4", "--save", path], threads, timeout=900)
        z = np.load(path)
        want = [z[k].copy() for k in ("h", "hu", "hv")]
    dt = float.fromhex(r["dt"])
    c = W.gauss_pulse(n)
    with swe_hip.HipSolver.from_case(c) as s:
        s.upload(c["h"], c["hu"], c["hv"], c["B"])
        s.step(dt, 2)
        s.upload(c["h"], c["hu"], c["hv"], c["B"])
        s.sync()
        t0 = time.perf_counter()
        s.step(dt, steps)
        el = time.perf_counter() - t0
        got = s.download()
    linf = 0.0
    for a, b in zip(got, want):
        scale = float(np.max(np.abs(b)))
        linf = max(linf, float(np.max(np.abs(a - b))) / (scale if scale > 0 else 1.0))
    bitwise = all(np.array_equal(a.view(np.uint64), b.view(np.uint64)) for a, b in zip(got, want))
    return {"workload": "gauss_pulse_256x256_g9.81_cfl0.4_1000_steps", "dt": dt,
            "cpu_kind": r["kind"], "cpu_threads": r["threads"],
            "cpu_gcell_per_s": round(n * n * steps / r["seconds"] / 1e9, 5),
            "gpu_gcell_per_s": round(n * n * steps / el / 1e9, 4),
            "max_rel_linf": linf, "bitwise": bool(bitwise)}


def ensemble_bench(steps, base_steps, check_every=32):
    """Config 4: 512 instances, sides log-uniform in [64, 1024], per-instance CFL 0.4 dt,
    advanced together (one tile work-list launch per stage) vs one context per instance.
    Two per-instance baselines: the launch loop (every stage, speed reduction and dt
    finalize of every instance launched on its own: 4 launches per instance and step,
    2048 per step) and the same kernels with each instance's steps replayed from a hipGraph.
    Rates in Gcell-updates/s."""
    import swe_hip
    from swe_hip import workloads as W

    sides = W.ensemble_sides(512)
    cases = [W.ensemble_member(i, int(n)) for i, n in enumerate(sides)]
    cells = int(sum(int(n) * int(n) for n in sides))
    b = swe_hip.HipBatch(cases)
    for i, c in enumerate(cases):
        b.upload(i, c["h"], c["hu"], c["hv"], c["B"])
    b.stable_dt(0.4)
    tile_rows, waves = b.autotune()
    # warmup with the timed step count: the dt history and the captured step graph are
    # sized and built before the timed region
    b.step_adaptive(0.4, steps)
    for i, c in enumerate(cases):
        b.upload(i, c["h"], c["hu"], c["hv"], c["B"])
    t0 = time.perf_counter()
    b.step_adaptive(0.4, steps)
    el_b = time.perf_counter() - t0
    rate_b = cells * steps / el_b / 1e9
    log(f"[bench] ensemble batched: {steps} steps {el_b:.3f} s -> {rate_b:.2f} Gcell/s "
        f"(tile_rows {tile_rows}, waves/CU {waves})")

    # per-instance launches on the same initial states, fewer steps (a rate)
    solvers = []
    for c in cases:
        s = swe_hip.HipSolver.from_case(c)
        s.upload(c["h"], c["hu"], c["hv"], c["B"])
        s.step_adaptive(0.4, base_steps)   # warmup: history sized, step graph captured
        s.upload(c["h"], c["hu"], c["hv"], c["B"])
        solvers.append(s)
    t0 = time.perf_counter()
    for s in solvers:
        s.step_adaptive(0.4, base_steps)
    el_s = time.perf_counter() - t0
    rate_s = cells * base_steps / el_s / 1e9
    log(f"[bench] ensemble per-instance (graph replay): {base_steps} steps {el_s:.3f} s -> "
        f"{rate_s:.3f} Gcell/s")

    # per-instance launch loop: the same kernels launched one by one
    for s, c in zip(solvers, cases):
        s.set_graph(False)
        s.upload(c["h"], c["hu"], c["hv"], c["B"])
        s.step_adaptive(0.4, base_steps)   # warmup with the timed step count
        s.upload(c["h"], c["hu"], c["hv"], c["B"])
    solvers[-1].sync()
    t0 = time.perf_counter()
    for s in solvers:
        s.step_adaptive(0.4, base_steps)
    el_l = time.perf_counter() - t0
    rate_l = cells * base_steps / el_l / 1e9
    log(f"[bench] ensemble per-instance (launch loop): {base_steps} steps {el_l:.3f} s -> "
        f"{rate_l:.3f} Gcell/s")

    # same number of steps in the batch from the same states; compare every check_every-th
    # instance bitwise
    for i, c in enumerate(cases):
        b.upload(i, c["h"], c["hu"], c["hv"], c["B"])
    b.step_adaptive(0.4, base_steps)
    same = True
    for i in range(0, len(cases), check_every):
        got = b.download(i)
        want = solvers[i].download()
        same &= all(np.array_equal(x.view(np.uint64), y.view(np.uint64)) for x, y in zip(got, want))
    for s in solvers:
        s.close()
    b.close()
    return {
        "workload": "512 instances, sides log-uniform [64,1024] (seed 4), 8-bump bathymetry, "
                    "Gaussian surface, reflecting, per-instance CFL 0.4",
        "cells": cells, "batched_steps": steps, "batched_gcell_per_s": round(rate_b, 3),
        "per_instance_steps": base_steps,
        "per_instance_launch_loop_gcell_per_s": round(rate_l, 4),
        "per_instance_graph_gcell_per_s": round(rate_s, 4),
        "speedup": round(rate_b / rate_l, 2), "speedup_vs_graph_replay": round(rate_b / rate_s, 2),
        "tile_rows": tile_rows, "waves_per_cu": waves,
        "bitwise_vs_per_instance": bool(same),
    }


def adjoint_bench(n=1024, steps=2000, plans=None):
    """Config 5: 1024^2, 2000 SSP-RK3 steps (fixed dt from CFL 0.4 at t = 0), gauge-misfit
    loss at 16 seeded cells.  Wall time of the plain forward (graph-replayed steps, state
    resident) vs the adjoint (forward with stored states + loss gradient + backward, state,
    loss gradient and gradients resident) for each storage plan: store every step
    ("auto" when it fits), the records-plus-checkpoints plan at 25% and 10% of that memory,
    and binomial checkpoints in 64 ... 2000 stored states."""
    import swe_hip
    from swe_hip import workloads as W

    case = W.adjoint_case(n, steps=steps)
    rows, cols = W.adjoint_gauges(n)
    with swe_hip.HipSolver.from_case(case) as s:
        s.upload(case["h"], case["hu"], case["hv"], case["B"])
        dt, _ = s.stable_dt(0.4)
        s.step(dt, 20)   # warmup (also captures the step driver's graph)
        s.upload(case["h"], case["hu"], case["hv"], case["B"])
        s.sync()
        t0 = time.perf_counter()
        s.step(dt, steps)
        t_fwd = time.perf_counter() - t0
        target = s.download()[0][rows, cols] * 0.99  # a synthetic target: 1% lower gauges
    res = {"workload": f"adjoint_{n}x{n}_{steps}_steps (16-bump bathymetry, surface bump, "
                       "background undulation + flow), gauge-misfit loss at 16 cells",
           "forward_s": round(t_fwd, 4)}
    arr = (n + 4) * (n + 4) * 8
    store_all = 9.0 * arr * steps   # every step with its two stage states
    runs = [("auto", 0, None), ("records_25pct", 0, 0.25 * store_all),
            ("records_10pct", 0, 0.10 * store_all)]
    runs += [(f"binomial_{k}", k, None) for k in (64, 128, 256, 512, steps)]
    if plans:
        runs = [r for r in runs if r[0] in plans]
    # the loss gradient and the gradients stay in device memory (torch tensors) inside the
    # timed region, as the forward's state does; host arrays only without torch
    try:
        import torch

        dev = torch.device("cuda:0") if torch.cuda.is_available() else None
    except Exception:  # pragma: no cover - torch absent
        dev = None
    if dev is not None:  # first use of torch's kernels (module loading) outside the timing
        w = torch.zeros((n + 4, n + 4), dtype=torch.float64, device=dev)
        ri, ci = torch.as_tensor(rows, device=dev), torch.as_tensor(cols, device=dev)
        w.index_put_((ri, ci), 2.0 * (w[ri, ci] - torch.as_tensor(target, device=dev)),
                     accumulate=True)
        torch.cuda.synchronize()
        del w
    ref_grad = None
    for label, nsnap, budget in runs:
        with swe_hip.HipSolver.from_case(case) as s:
            s.upload(case["h"], case["hu"], case["hv"], case["B"])
            with swe_hip.HipAdjoint(s, steps, nsnap, budget_bytes=budget) as adj:
                if dev is not None:
                    shape = s.shape
                    hN_d, huN_d, hvN_d = (torch.empty(shape, dtype=torch.float64, device=dev)
                                          for _ in range(3))
                    lam_d = torch.zeros(shape, dtype=torch.float64, device=dev)
                    z_d = torch.zeros(shape, dtype=torch.float64, device=dev)
                    g_d = [torch.empty(shape, dtype=torch.float64, device=dev) for _ in range(4)]
                    r_d = torch.as_tensor(rows, device=dev)
                    c_d = torch.as_tensor(cols, device=dev)
                    tgt_d = torch.as_tensor(target, device=dev)
                    torch.cuda.synchronize()
                    t0 = time.perf_counter()
                    adj.forward(dt=dt)
                    s.download_device(hN_d, huN_d, hvN_d)
                    lam_d.index_put_((r_d, c_d), 2.0 * (hN_d[r_d, c_d] - tgt_d), accumulate=True)
                    torch.cuda.synchronize()
                    adj.backward_device(lam_d, z_d, z_d, *g_d)
                    torch.cuda.synchronize()
                    t_adj = time.perf_counter() - t0
                    g = [x.cpu().numpy() for x in g_d]
                else:
                    t0 = time.perf_counter()
                    adj.forward(dt=dt)
                    hN = s.download()[0]
                    lam = np.zeros_like(hN)
                    np.add.at(lam, (rows, cols), 2.0 * (hN[rows, cols] - target))
                    z = np.zeros_like(hN)
                    g = adj.backward(lam, z, z)
                    t_adj = time.perf_counter() - t0
                info = adj.info()
                plan = adj.plan()
        res[label] = {"adjoint_s": round(t_adj, 4), "ratio": round(t_adj / t_fwd, 3),
                      "nsnap": info["nsnap"], "stage_records": info["records"],
                      "record_steps": plan["record_steps"], "ckpt_every": plan["ckpt_every"],
                      "store_gb": round(plan["store_bytes"] / 1e9, 3),
                      "store_frac_of_store_all": round(plan["store_bytes"] / store_all, 4),
                      "recomputed_steps": round(plan["recomputed_steps"], 1),
                      "forward_ms": round(info["forward_ms"], 1),
                      "backward_ms": round(info["backward_ms"], 1),
                      "grad_finite": bool(np.isfinite(g[0]).all() and np.isfinite(g[3]).all())}
        if ref_grad is None:
            ref_grad = g
        res[label]["grad_bitwise_vs_auto"] = all(
            np.array_equal(a.view(np.uint64), b.view(np.uint64)) for a, b in zip(g, ref_grad))
        log(f"[bench] adjoint {label}: {t_adj:.3f} s vs forward {t_fwd:.3f} s "
            f"-> {t_adj / t_fwd:.2f}x ({info}, {plan})")
    res["ratio"] = res["auto"]["ratio"] if "auto" in res else None
    # the least stored memory among the runs that stay within 4x the forward time
    ok = [(v["store_frac_of_store_all"], k) for k, v in res.items()
          if isinstance(v, dict) and v.get("ratio", 99) <= 4.0]
    res["least_memory_within_4x"] = min(ok)[1] if ok else None
    return res


def serial_bench(n=4096, steps=20, cpu_n=1024, cpu_steps=2):
    """The modular serial scheme family (SWE_SCHEME_SERIAL*, the reference's SerialSolver
    combinations) on an n^2 wet workload with bathymetry and flow (the adjoint workload's
    generator; the plain schemes ignore B).  Per combination: fixed-dt Gcell-updates/s from
    HIP events around every stage launch, the stage-2 kernel's algorithmic bandwidth
    (plain: 72 B/cell, hydrostatic: 80 B/cell), and the reference's own modular code
    (oracle/_ref, single-threaded like the reference Serial backend) on a cpu_n^2 sample of
    the same generator, checked bitwise against the GPU on that sample."""
    import swe_hip
    from swe_hip import workloads as W
    import oracle

    combos = [(0, r, q) for r in (0, 1) for q in (0, 1, 2)] + [(1, 0, 0), (1, 1, 0)]
    rn = {0: "MUSCL", 1: "PiecewiseConst"}
    qn = {0: "HLL", 1: "Rusanov", 2: "ROE"}
    case = W.adjoint_case(n, steps=steps)
    small = W.adjoint_case(cpu_n, steps=cpu_steps)
    use_ref = oracle.ref_modular() is not None
    rows = []
    for bathy, recon, rie in combos:
        name = f"{'hydrostatic' if bathy else 'plain'} {rn[recon]} + {qn[rie]}"
        with swe_hip.HipSolver.from_case(case) as s:
            s.set_scheme(swe_hip.SCHEME_SERIAL if bathy else swe_hip.SCHEME_SERIAL_PLAIN,
                         recon, rie)
            s.upload(case["h"], case["hu"], case["hv"], case["B"])
            dt = 0.2 * min(case["dx"], case["dy"])
            s.step(dt, 2)
            st, _ = s.time_steps(dt, steps)
        cells = n * n
        rate = cells / (sum(st) * 1e-3) / 1e9
        b2 = (80 if bathy else 72) * cells
        # CPU: the reference's modular code on the small sample, and bitwise check
        gs = [small[k].copy() for k in ("h", "hu", "hv")]
        dts = 0.2 * min(small["dx"], small["dy"])
        with swe_hip.HipSolver.from_case(small) as s:
            s.set_scheme(swe_hip.SCHEME_SERIAL if bathy else swe_hip.SCHEME_SERIAL_PLAIN,
                         recon, rie)
            s.upload(small["h"], small["hu"], small["hv"], small["B"])
            s.step(dts, cpu_steps)
            got = s.download()
        cs = [small[k].copy() for k in ("h", "hu", "hv")]
        t0 = time.perf_counter()
        if use_ref:
            rc = oracle.ref_mod_run(cpu_n, cpu_n, small["lx"], small["ly"], bathy, recon, rie,
                                    *cs, small["B"], cpu_steps, dts)
        else:
            gr = oracle.grid(cpu_n, cpu_n, lx=small["lx"], ly=small["ly"])
            rc = oracle.mod_run(gr, bathy, recon, rie, *cs, small["B"], cpu_steps, dts)
        el = time.perf_counter() - t0
        same = rc == 0 and all(np.array_equal(a.view(np.uint64), b.view(np.uint64))
                               for a, b in zip(got, cs))
        rows.append({"scheme": name, "gcell_per_s": round(rate, 3),
                     "stage_ms": [round(x, 4) for x in st],
                     "stage2_gbs": round(b2 / (st[1] * 1e-3) / 1e9, 1),
                     "stage2_frac_hbm": round(b2 / (st[1] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "cpu_gcell_per_s": round(cpu_n * cpu_n * cpu_steps / el / 1e9, 5),
                     "bitwise_vs_cpu": bool(same)})
        log(f"[bench] serial scheme {name}: {rate:.2f} Gcell/s, CPU "
            f"{rows[-1]['cpu_gcell_per_s']} Gcell/s, bitwise {same}")
    return {"workload": f"{n}x{n} wet, 16-bump bathymetry + background flow, reflecting, "
                        f"fixed dt, {steps} steps (HIP events per stage)",
            "cpu": {"kind": "reference" if use_ref else "port", "cores": 1 if use_ref else
                    oracle.lib().oracle_num_threads(),
                    "sample": f"{cpu_n}x{cpu_n}, {cpu_steps} steps of the same generator"},
            "schemes": rows}


def fused_bench(case, dt, three_launch_value, rows_list=(64, 128), steps=100):
    """The fused one-launch step (temporal blocking across the three stages) on the headline
    workload, per tile height: Gcell-updates/s, its ratio to the three-launch step, and the
    final state after warmup + steps against the three-launch path's (bitwise)."""
    import swe_hip

    cells = case["nx"] * case["ny"]
    with swe_hip.HipSolver.from_case(case) as ref:
        ref.upload(case["h"], case["hu"], case["hv"], case["B"])
        ref.step(dt, 2 + steps)
        want = ref.download()
    out = {"workload": "the headline dam break, one launch per step", "steps": steps,
           "algorithmic_bytes_per_cell_update": 56, "three_launch_gcell_per_s": three_launch_value,
           "tiles": []}
    for rows in rows_list:
        with swe_hip.HipSolver.from_case(case) as s:
            s.upload(case["h"], case["hu"], case["hv"], case["B"])
            s.step_fusion = rows
            s.step(dt, 2)
            s.sync()
            t0 = time.perf_counter()
            s.step(dt, steps)
            el = time.perf_counter() - t0
            got = s.download()
        v = cells * steps / el / 1e9
        same = all(np.array_equal(a.view(np.uint64), b.view(np.uint64)) for a, b in zip(got, want))
        rec = {"rows": rows, "gcell_per_s": round(v, 3), "ms_per_step": round(el / steps * 1e3, 4),
               "vs_three_launch": round(v / three_launch_value, 4),
               "hbm_frac_at_56B": round(v * 1e9 * 56 / (HBM_PEAK_GBS * 1e9), 4),
               "bitwise_vs_three_launch": bool(same)}
        log(f"[bench] fused step, {rows}-row tiles: {rec}")
        out["tiles"].append(rec)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile-rows", type=int, default=0)
    ap.add_argument("--stage-tile-rows", type=str, default="",
                    help="tile heights of stages 1, 2, 3 (comma-separated)")
    ap.add_argument("--event-steps", type=int, default=20)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--ensemble-base-steps", type=int, default=32,
                    help="per-instance baseline steps (32: one replay of the step driver's graph)")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-serial", action="store_true")
    ap.add_argument("--no-fused", action="store_true")
    ap.add_argument("--fused-rows", type=str, default="64,128",
                    help="tile heights of the fused one-launch step to time")
    ap.add_argument("--serial-n", type=int, default=4096)
    ap.add_argument("--adjoint-steps", type=int, default=2000)
    ap.add_argument("--adjoint-plans", type=str, default="",
                    help="comma-separated subset of the adjoint storage plans to run")
    ap.add_argument("--workload", choices=("dam_break", "wet_bathymetry"), default="dam_break",
                    help="dam_break: the headline config (flat bottom); wet_bathymetry: the "
                         "same size with 16 Gaussian bumps and background flow (sloped faces)")
    ap.add_argument("--precision", choices=("exact", "fast"), default="exact",
                    help="fast: SWE_PRECISION_FAST (not bitwise; never the headline)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the slab decomposition across GPUs is not "
              "implemented. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import swe_hip
    from swe_hip import workloads as W

    if swe_hip.device_count() < 1:
        raise SystemExit("no HIP device visible")

    n = args.n
    log(f"[bench] generating {args.workload} {n}^2")
    case = W.dam_break(n) if args.workload == "dam_break" else W.adjoint_case(n)
    s = swe_hip.HipSolver.from_case(case)
    if args.tile_rows:
        s.tile_rows = args.tile_rows
    if args.stage_tile_rows:
        s.stage_tile_rows = [int(x) for x in args.stage_tile_rows.split(",")]
    if args.precision == "fast":
        s.precision = swe_hip.PRECISION_FAST
    s.upload(case["h"], case["hu"], case["hv"], case["B"])
    dt, max_speed = s.stable_dt(0.45 if args.workload == "dam_break" else 0.4)
    m0 = mass(case["h"], case["ng"])
    log(f"[bench] dt={dt!r} max_speed={max_speed!r} stage tile rows={s.stage_tile_rows}")

    # warmup (also builds the step graph)
    s.step(dt, max(args.warmup, 1))
    s.sync()
    t0 = time.perf_counter()
    s.step(dt, args.steps)        # returns after a stream sync + non-finite check
    el = time.perf_counter() - t0
    cells = n * n
    value = cells * args.steps / el / 1e9
    ms_step = el / args.steps * 1e3
    log(f"[bench] {args.steps} steps in {el:.3f} s: {ms_step:.3f} ms/step, {value:.2f} Gcell/s")

    # per-stage event timing (direct launches, events on the solver stream)
    stage_ms, ev_total = s.time_steps(dt, args.event_steps)
    log(f"[bench] stage ms {stage_ms}, event total {ev_total:.3f} ms / {args.event_steps} steps")
    flat_b = s.flat_bottom
    bpc = BYTES_PER_CELL_STAGE_FLAT_B if flat_b else BYTES_PER_CELL_STAGE
    general = None
    steps_done = max(args.warmup, 1) + args.steps + args.event_steps
    if flat_b:
        # the same grid through the kernels that read B (the general path), for comparison
        s.set_flat_bottom(False)
        gsteps = max(1, min(args.steps, 300))
        s.step(dt, 2)
        s.sync()
        t0 = time.perf_counter()
        s.step(dt, gsteps)
        gel = time.perf_counter() - t0
        gev = max(1, min(args.event_steps, 50))
        g_ms, _ = s.time_steps(dt, gev)
        s.set_flat_bottom(True)
        steps_done += 2 + gsteps + gev
        general = {"gcell_per_s": round(cells * gsteps / gel / 1e9, 4), "steps": gsteps,
                   "ms_per_step": round(gel / gsteps * 1e3, 4),
                   "stage_ms": [round(x, 4) for x in g_ms],
                   "bytes_per_cell_update": sum(BYTES_PER_CELL_STAGE)}
        log(f"[bench] general (B read) path: {general['gcell_per_s']:.2f} Gcell/s, "
            f"stage ms {g_ms}")
    h_end = s.download()[0]
    m1 = mass(h_end, case["ng"])
    stage_bytes = [b * cells for b in bpc]
    dom = 1  # stage 2 kernel: 80 (72) B/cell, the largest per launch
    ach_dom = stage_bytes[dom] / (stage_ms[dom] * 1e-3) / 1e9
    step_bytes = sum(stage_bytes)
    ach_step = step_bytes / (sum(stage_ms) * 1e-3) / 1e9
    traffic, valu = None, None
    tf = os.path.join(REPO, "profiles", "stage_kernel_dam_break_8192_traffic.json")
    if os.path.exists(tf) and n == 8192 and args.workload == "dam_break" and \
            args.precision == "exact":
        prof = json.load(open(tf))
        traffic = prof.get("stage2_hbm_bytes_per_launch")
        qc = prof.get("stage2_valu_active_quad_cycles_per_launch")
        if qc:
            # the VALU issue floor: the stage-2 kernel's VALU quad-cycles (SQ_ACTIVE_INST_VALU,
            # transcendental passes included) spread over 256 CUs x 4 SIMDs at the rated clock
            floor_ms = qc * 4 / (256 * 4) / (SHADER_CLOCK_GHZ * 1e9) * 1e3
            valu = {"active_quad_cycles_per_launch": qc,
                    "insts_per_launch": prof.get("stage2_valu_insts_per_launch"),
                    "issue_floor_ms": round(floor_ms, 4), "clock_ghz": SHADER_CLOCK_GHZ,
                    "issue_frac": round(floor_ms / stage_ms[dom], 4),
                    "valu_busy_frac_counters": prof.get("stage2_valu_busy_frac"),
                    "source": "profiles/stage_kernel_dam_break_8192_pmc.md"}

    out = {
        "metric": "FP64 Gcell-updates/s @8192^2 (dam break, SSP-RK3 MUSCL+HLL+HR)",
        "value": round(value, 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded dam break)",
        "config": {"workload": (f"dam_break_{n}x{n}_fp64_reflecting_cfl0.45"
                                if args.workload == "dam_break" else
                                f"wet_bathymetry_{n}x{n}_fp64_reflecting_cfl0.4"),
                   "nx": n, "ny": n, "ng": 2, "dt": dt, "parallelism": "single GPU",
                   "tile_rows": list(s.stage_tile_rows), "precision": args.precision,
                   "bottom_not_read": bool(flat_b)},
        "roofline": {
            "bound": "hbm",
            "kernel": (f"stage_kernel<2> (SSP-RK3 stage 2, {bpc[1]} B/cell"
                       + (": bottom +0.0 everywhere, not read)" if flat_b else ")")),
            "achieved": round(ach_dom, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach_dom / HBM_PEAK_GBS, 4), "traffic": traffic,
            # the roof that binds: FP64 VALU issue (the HBM fraction above is the metric's
            # pricing; the kernel is limited by its bitwise FP64 arithmetic, not by bytes)
            "binding_roof": "fp64_valu_issue" if valu else None,
            "valu": valu,
        },
        "roofline_step": {
            "bytes_per_cell_update": sum(bpc), "achieved": round(ach_step, 1),
            "frac": round(ach_step / HBM_PEAK_GBS, 4), "stage_ms": [round(x, 4) for x in stage_ms],
        },
        "pct_hbm_peak": round(100.0 * value * 1e9 * sum(bpc) / (HBM_PEAK_GBS * 1e9), 2),
        "general_bottom_path": general,
        "mass_rel_change": abs(m1 - m0) / m0,
        "steps_total_on_state": steps_done,
    }
    s.close()

    if not args.no_cpu:
        from oracle import cpu_baseline as cb

        log(f"[bench] CPU baseline: {args.cpu_steps} steps")
        info, ref_sha = cpu_baseline(case, args.cpu_steps, dt, args.workload)
        # same steps on the GPU, compared bitwise (sha256 of the full arrays)
        with swe_hip.HipSolver.from_case(case) as s2:
            s2.upload(case["h"], case["hu"], case["hv"], case["B"])
            s2.step(dt, args.cpu_steps)
            got = s2.download()
        same = cb.digest(*got) == ref_sha
        out["cpu_parity_bitwise"] = bool(same) if args.precision == "exact" else None
        log(f"[bench] cpu {info['value']:.4f} Gcell/s ({info['cores']} threads, "
            f"{info['binding']}, host {info['host']}), bitwise={same}")
        if not args.no_c1:
            info["c1"] = c1_record(info["threads"])
            log(f"[bench] C1: {info['c1']}")
        out["cpu_baseline"] = info
    if not args.no_fused and args.precision == "exact":
        out["fused_step"] = fused_bench(case, dt, out["value"],
                                        [int(r) for r in args.fused_rows.split(",") if r])
    if not args.no_ensemble:
        out["batched"] = ensemble_bench(args.ensemble_steps, args.ensemble_base_steps)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_bench(steps=args.adjoint_steps,
                                       plans=[p for p in args.adjoint_plans.split(",") if p])
    if not args.no_serial:
        out["serial_schemes"] = serial_bench(n=args.serial_n)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
