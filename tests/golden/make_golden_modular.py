"""Write tests/golden/modular.npz from the reference's own modular serial solver parts.

Run in a container that has the reference tree (it is not needed afterwards):

    make -C oracle ref            # builds oracle/_ref/libswe_ref_modular.so
    python tests/golden/make_golden_modular.py

Every expected output comes from oracle/_ref/libswe_ref_modular.so, i.e. the unmodified
reference headers (include/backend/serial/{reconstruction,riemann,flux_assembly,
finite_volume,ssp_rk3}, include/boundary/reflecting_walls.hpp,
include/solver_assembly/time_utils.hpp) compiled here and driven by the step loop of
include/backend/serial/serial_solver.hpp:180-300 (g = 9.80665).  Inputs are stored with
the outputs.

Arrays (np.savez_compressed; load with numpy's default allow_pickle=False):
  riemann_L, riemann_R (n, 3)       scalar states (random + edge cases)
  riemann_{x,y}_{hll,rusanov,roe}   fluxes (n, 3); riemann_*_fail: 1 where HLL throws
  grid                              nx, ny, lx, ly
  in_{h,hu,hv,B}_{wet,dry}          ghost-filled inputs (wet: all h > 0.5; dry: some cells
                                    dry or at eps)
  rhs_<case>_<bathy>_<recon>_<riemann>   (3, nxt, nyt) rhs; rc where the reference throws
  run_<case>_<bathy>_<recon>_<riemann>   (3, nxt, nyt) state after `steps` steps of dt
  dt_<case>                          the CFL-check value of time_utils.hpp at cfl 0.4
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

import oracle  # noqa: E402

NAMES_RIEMANN = {oracle.RIEMANN_HLL: "hll", oracle.RIEMANN_RUSANOV: "rusanov",
                 oracle.RIEMANN_ROE: "roe"}
NAMES_RECON = {oracle.RECON_MUSCL: "muscl", oracle.RECON_PC: "pc"}
NX, NY, LX, LY = 14, 11, 1.75, 1.375
STEPS = 5


def scalar_states(rng):
    n = 600
    L = np.stack([0.5 + 1.5 * rng.random(n), rng.normal(0, 1.5, n), rng.normal(0, 1.5, n)], 1)
    R = np.stack([0.5 + 1.5 * rng.random(n), rng.normal(0, 1.5, n), rng.normal(0, 1.5, n)], 1)
    edge = [
        ((1.0, 0.0, 0.0), (1.0, 0.0, 0.0)),          # equal states at rest
        ((1.0, 0.3, -0.2), (1.0, 0.3, -0.2)),        # equal moving states
        ((0.0, 0.0, 0.0), (1.0, 0.0, 0.0)),          # dry | wet
        ((1.0, 0.0, 0.0), (0.0, 0.0, 0.0)),          # wet | dry
        ((0.0, 0.0, 0.0), (0.0, 0.0, 0.0)),          # dry | dry (ROE divides by 0)
        ((1e-8, 1e-9, 0.0), (1.0, 0.1, 0.0)),        # h == eps on the left
        ((2e-8, 1e-9, 0.0), (1e-8, -1e-9, 0.0)),     # eps neighbourhood
        ((1.0, -0.0, 0.0), (1.0, 0.0, -0.0)),        # signed zeros
        ((1.0, 20.0, 0.0), (1.0, 20.0, 0.0)),        # supercritical to +x
        ((1.0, -20.0, 0.0), (1.0, -20.0, 0.0)),      # supercritical to -x
        ((1.0, 0.0, 20.0), (1.0, 0.0, 20.0)),        # supercritical to +y
        ((1.0, 0.0, -20.0), (0.5, 0.0, -20.0)),      # supercritical to -y
        ((float("nan"), 0.0, 0.0), (1.0, 0.0, 0.0)),  # HLL throws
    ]
    L = np.concatenate([L, np.array([e[0] for e in edge])])
    R = np.concatenate([R, np.array([e[1] for e in edge])])
    return np.ascontiguousarray(L), np.ascontiguousarray(R)


def case_inputs(rng, dry: bool):
    gr = oracle.grid(NX, NY, lx=LX, ly=LY)
    shape = (NX + 4, NY + 4)
    B = 0.4 * rng.random(shape)
    if dry:
        h = 0.8 * rng.random(shape)
        h[rng.random(shape) < 0.25] = 0.0
        h[rng.random(shape) < 0.05] = 1e-8
    else:
        h = 0.6 + 0.6 * rng.random(shape)
    hu = 0.3 * rng.standard_normal(shape)
    hv = 0.3 * rng.standard_normal(shape)
    hu[h == 0.0] = 0.0
    hv[h == 0.0] = 0.0
    oracle.mod_apply_bc(gr, B)
    oracle.mod_apply_bc(gr, h, hu, hv)
    return [np.ascontiguousarray(a) for a in (h, hu, hv, B)]


def main():
    assert oracle.ref_modular() is not None, "build the reference harness first (make -C oracle ref)"
    rng = np.random.default_rng(2024)
    out = {"grid": np.array([NX, NY, LX, LY]), "steps": np.array(STEPS)}
    L, R = scalar_states(rng)
    out["riemann_L"], out["riemann_R"] = L, R
    for d in ("x", "y"):
        for kind, nm in NAMES_RIEMANN.items():
            F = np.zeros_like(L)
            fail = np.zeros(len(L), np.int32)
            for i in range(len(L)):
                F[i], fail[i] = oracle.ref_mod_riemann(d, kind, L[i], R[i])
            out[f"riemann_{d}_{nm}"] = F
            out[f"riemann_{d}_{nm}_fail"] = fail
    for case in ("wet", "dry"):
        h, hu, hv, B = case_inputs(rng, case == "dry")
        for k, a in zip(("h", "hu", "hv", "B"), (h, hu, hv, B)):
            out[f"in_{k}_{case}"] = a
        dt = oracle.ref_mod_stable_dt(NX, NY, LX, LY, h, hu, hv, 0.4)
        out[f"dt_{case}"] = np.array(dt)
        step_dt = 0.2 * LX / NX / 4.0
        out[f"stepdt_{case}"] = np.array(step_dt)
        for bathy in (0, 1):
            for recon, rn in NAMES_RECON.items():
                for rie, qn in NAMES_RIEMANN.items():
                    tag = f"{case}_{bathy}_{rn}_{qn}"
                    s = [a.copy() for a in (h, hu, hv)]
                    rc, r = oracle.ref_mod_rhs(NX, NY, LX, LY, bathy, recon, rie, *s, B)
                    out[f"rhs_{tag}"] = np.stack(r)
                    out[f"rhsrc_{tag}"] = np.array(rc)
                    s = [a.copy() for a in (h, hu, hv)]
                    rc = oracle.ref_mod_run(NX, NY, LX, LY, bathy, recon, rie, *s, B, STEPS,
                                            step_dt)
                    out[f"run_{tag}"] = np.stack(s)
                    out[f"runrc_{tag}"] = np.array(rc)
    np.savez_compressed(os.path.join(HERE, "modular.npz"), **out)
    print("wrote", os.path.join(HERE, "modular.npz"))


if __name__ == "__main__":
    main()
