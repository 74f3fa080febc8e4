"""Golden initial states of the host application's generators, written from the
reference's own setup code.

Run in the build container (needs oracle/_ref, built from the reference tree by
``make -C oracle ref``):

    python tests/golden/make_golden_setup.py

For every initial-condition x bathymetry type of the reference's configuration schema
(configs/config.hpp:11-13) this calls oracle.ref_initial_state -- the reference's
apply_initial_condition.hpp:13-42 and apply_bathymetry.hpp:12-37 compiled where they lie,
followed by the solver constructor's ghost fills (omp_solver.hpp:118-127) -- and stores the
arrays in tests/golden/setup_states.npz.  Two parameter sets: the one of
tests/test_host_app.py (40 x 32 cells on 10 x 8) and the TOML forms of config 1
(GaussInitial, sigma = sqrt(0.005), on the unit square, 64^2 here) and config 3 (DamBreak
h0 = 2, dam_height = 1, dam_x = 10 on [0, 20]^2, 48^2 here).
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

import oracle  # noqa: E402

BASE_MESH = dict(nx=40, ny=32, ng=2, lx=10.0, ly=8.0)
BASE_IC = dict(h0=1.0, peak_height=0.5, x0=6.0, y0=4.0, sigma_x=1.0, sigma_y=0.5,
               dam_height=2.5, dam_x=5.0, dam_x0=5.0, dam_y0=4.0, dam_radius=2.0)
BASE_BT = dict(b0=0.1, bathy_x0=4.0, bathy_y0=3.0, bathy_sigma_x=1.5, bathy_sigma_y=0.75,
               bathy_peak_height=0.4)


def cases():
    for ic in oracle.IC_TYPES:
        for bt in oracle.BATHY_TYPES:
            yield f"{ic}__{bt}", BASE_MESH, ic, BASE_IC, bt, BASE_BT
    s = float(np.sqrt(0.005))
    yield ("c1_gauss__Flat", dict(nx=64, ny=64, ng=2, lx=1.0, ly=1.0), "GaussInitial",
           dict(h0=1.0, peak_height=0.5, x0=0.5, y0=0.5, sigma_x=s, sigma_y=s), "Flat",
           dict(b0=0.0))
    yield ("c3_dam_break__Flat", dict(nx=48, ny=48, ng=2, lx=20.0, ly=20.0), "DamBreak",
           dict(h0=2.0, dam_height=1.0, dam_x=10.0), "Flat", dict(b0=0.0))


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libswe_ref.so is missing: make -C oracle ref")
    arrays, manifest = {}, {}
    for name, mesh, ic, icp, bt, btp in cases():
        h, hu, hv, B = oracle.ref_initial_state(mesh["nx"], mesh["ny"], mesh["ng"], mesh["lx"],
                                                mesh["ly"], ic, icp, bt, btp)
        for k, a in zip(("h", "hu", "hv", "B"), (h, hu, hv, B)):
            arrays[f"{name}__{k}"] = a
        manifest[name] = {"mesh": mesh, "initial_condition": {"type": ic, **icp},
                          "bathymetry": {"type": bt, **btp}}
    np.savez_compressed(os.path.join(HERE, "setup_states.npz"), **arrays)
    with open(os.path.join(HERE, "setup_states.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    print(f"wrote {len(manifest)} setups")


if __name__ == "__main__":
    main()
