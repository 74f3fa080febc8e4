"""CPU baseline runner: the reference's compiled OpenMP kernels timed on the host.

TEST INFRASTRUCTURE ONLY (the timed CPU baseline of bench.py).  Run as a child process so
that the OpenMP runtime starts with the thread count and binding chosen here (libgomp
reads OMP_* when it loads), the way the reference's job script starts the OpenMP solver
(scripts/leo3e/run_openmp.sh:111-114: OMP_NUM_THREADS = cores, OMP_PROC_BIND=true,
OMP_PLACES=cores):

    python3 -m oracle.cpu_baseline --workload dam_break --n 8192 --steps 10 --dt <hex>
    python3 -m oracle.cpu_baseline --workload gauss_pulse --n 256 --steps 1000 --cfl 0.4 \
        --save <file.npz>

The child builds the seeded workload (swe_hip.workloads, numpy only), applies the
bathymetry ghost fill, runs the reference's rk3_step loop through oracle/_ref (or the
oracle port when the reference harness was not built), times only that loop, and prints
one JSON line: seconds, the dt used, the sha256 of the final h, hu, hv arrays (ghosts
included) and, with --save, writes those arrays for an exact comparison.

host_cpus() reports what the host offers: physical cores (lscpu), the CPUs of this
process's affinity mask, the cores among them and the cgroup CPU quota.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import math
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)


def _lscpu_cores():
    """{cpu: (socket, core)} from `lscpu -p=CPU,CORE,SOCKET`, or None."""
    try:
        out = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True,
                             check=True).stdout
    except Exception:
        return None
    m = {}
    for line in out.splitlines():
        if not line or line.startswith("#"):
            continue
        cpu, core, sock = (int(x) if x else 0 for x in line.split(","))
        m[cpu] = (sock, core)
    return m or None


def _cgroup_quota():
    """CPUs allowed by the cgroup v2 (cpu.max) or v1 (cfs quota/period) limit, or None."""
    try:
        q, p = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            return int(q) / int(p)
    except Exception:
        pass
    try:
        q = int(open("/sys/fs/cgroup/cpu/cpu.cfs_quota_us").read())
        p = int(open("/sys/fs/cgroup/cpu/cpu.cfs_period_us").read())
        if q > 0:
            return q / p
    except Exception:
        pass
    return None


def host_cpus():
    """Physical cores of the host, CPUs and cores of this process's affinity mask, the
    cgroup quota, and the OpenMP thread count the baseline uses: one thread per physical
    core available to the process, capped by the quota and by OMP_NUM_THREADS if set."""
    cores = _lscpu_cores()
    aff = sorted(os.sched_getaffinity(0))
    quota = _cgroup_quota()
    phys = len(set(cores.values())) if cores else None
    aff_cores = len({cores.get(c, (0, c)) for c in aff}) if cores else len(aff)
    threads = aff_cores
    if quota is not None:
        threads = min(threads, max(1, math.floor(quota)))
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        threads = min(threads, int(env))
    return {"physical_cores": phys, "affinity_cpus": len(aff), "affinity_cores": aff_cores,
            "cgroup_quota_cpus": quota, "threads": max(1, threads)}


def child_env(threads: int) -> dict:
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(threads)
    env["OMP_PROC_BIND"] = "true"
    env["OMP_PLACES"] = "cores"
    return env


def run_child(args: list[str], threads: int, timeout: float = 600.0) -> dict:
    """Start this module as a child with the OpenMP environment above; parse its JSON."""
    res = subprocess.run([sys.executable, "-m", "oracle.cpu_baseline", *args], cwd=REPO,
                         env=child_env(threads), capture_output=True, text=True,
                         timeout=timeout)
    if res.returncode != 0:
        raise RuntimeError(f"cpu baseline child failed ({res.returncode}): {res.stderr[-2000:]}")
    return json.loads(res.stdout.strip().splitlines()[-1])


def digest(*arrays) -> list[str]:
    import numpy as np

    return [hashlib.sha256(np.ascontiguousarray(a).view(np.uint8)).hexdigest() for a in arrays]


def _main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=("dam_break", "gauss_pulse", "wet_bathymetry"),
                    required=True)
    ap.add_argument("--n", type=int, required=True)
    ap.add_argument("--steps", type=int, required=True)
    ap.add_argument("--dt", type=str, default=None, help="fixed dt as float.hex()")
    ap.add_argument("--cfl", type=float, default=None, help="dt from this CFL at t = 0")
    ap.add_argument("--g", type=float, default=None)
    ap.add_argument("--save", type=str, default=None)
    a = ap.parse_args()

    sys.path.insert(0, REPO)
    sys.path.insert(0, os.path.join(REPO, "shallow-water_amd"))
    import numpy as np

    import oracle
    from swe_hip import workloads as W

    if a.workload == "dam_break":
        c = W.dam_break(a.n) if a.g is None else W.dam_break(a.n, g=a.g)
    elif a.workload == "wet_bathymetry":
        c = W.adjoint_case(a.n) if a.g is None else W.adjoint_case(a.n, g=a.g)
    else:
        c = W.gauss_pulse(a.n) if a.g is None else W.gauss_pulse(a.n, g=a.g)
    gr = oracle.grid(c["nx"], c["ny"], c["ng"], c["dx"], c["dy"], g=c["g"], bc=c["bc"])
    h, hu, hv = (c[k].copy() for k in ("h", "hu", "hv"))
    B = c["B"].copy()
    oracle.apply_bc(gr, B)
    R = oracle.ref_for_g(c["g"]) if c["bc"] == 0 else None
    if a.dt is not None:
        dt = float.fromhex(a.dt)
    else:
        dt = (R.ref_stable_dt(gr.nx, gr.ny, gr.ng, gr.dx, gr.dy, oracle._p(h), oracle._p(hu),
                              oracle._p(hv), a.cfl) if R is not None
              else oracle.stable_dt(gr, h, hu, hv, a.cfl)[0])
    t0 = time.perf_counter()
    if R is not None:
        kind = "reference"
        threads = R.ref_num_threads()
        rc, _ = oracle.ref_run(gr, h, hu, hv, B, a.steps, dt=dt, lib=R)
    else:
        kind = "port"
        threads = oracle.lib().oracle_num_threads()
        rc, _, _ = oracle.run(gr, h, hu, hv, B, a.steps, dt=dt)
    el = time.perf_counter() - t0
    if a.save:
        np.savez(a.save, h=h, hu=hu, hv=hv)
    print(json.dumps({"kind": kind, "threads": threads, "rc": int(rc), "seconds": el,
                      "dt": dt.hex(), "g": c["g"], "cells": c["nx"] * c["ny"],
                      "steps": a.steps, "sha256": digest(h, hu, hv),
                      "bind": os.environ.get("OMP_PROC_BIND"),
                      "places": os.environ.get("OMP_PLACES")}))


if __name__ == "__main__":
    _main()
