"""The oracle against the golden fixtures written from the reference's compiled kernels.

Bar: bitwise.  These pin the C restatement (oracle/swe_oracle.c) to the reference's own
include/backend/openmp/omp_kernels.hpp + the rk3_step sequence of omp_solver.hpp:250-338,
without needing the reference tree (the fixtures are committed)."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
from swe_hip import workloads as W

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MANIFEST = json.load(open(os.path.join(GOLDEN, "manifest.json")))


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def same(a, b):
    return np.array_equal(bits(np.asarray(a, dtype=np.float64)), bits(np.asarray(b, dtype=np.float64)))


def load(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def test_manifest_g_is_reference():
    assert MANIFEST["g"] == 9.80665


def test_minmod_known_answers():
    z = load("scalar")
    got = np.array([oracle.minmod(a, b) for a, b in zip(z["minmod_a"], z["minmod_b"])])
    assert same(got, z["minmod"])


@pytest.mark.parametrize("direction", ["x", "y"])
def test_hll_known_answers(direction):
    z = load("scalar")
    want = z["hll_" + direction]
    got = np.array([oracle.hll(direction, *row) for row in z["hll_in"]])
    assert same(got, want)
    # the fixture exercises every branch of hll_x/hll_y (omp_kernels.hpp:72-85)
    hL, huL, hR, huR = z["hll_in"][:, 0], z["hll_in"][:, 1], z["hll_in"][:, 3], z["hll_in"][:, 4]
    assert (hL == 0).any() and (hR == 0).any()


@pytest.mark.parametrize("name", ["rhs_random_17x13", "rhs_wetdry_32x32", "rhs_lake_24"])
def test_faces_and_rhs(name):
    m, z = MANIFEST[name], load(name)
    gr = oracle.grid(m["nx"], m["ny"], m["ng"], m["dx"], m["dy"])
    f = oracle.faces(gr, z["h"], z["hu"], z["hv"], z["B"])
    keys = ["fxm_h", "fxm_hu", "fxm_hv", "fxp_h", "fxp_hu", "fxp_hv",
            "fym_h", "fym_hu", "fym_hv", "fyp_h", "fyp_hu", "fyp_hv"]
    for k, a in zip(keys, f):
        assert same(a, z[k]), k
    r = oracle.rhs(gr, z["h"].copy(), z["hu"].copy(), z["hv"].copy(), z["B"])
    for k, a in zip(("rhs_h", "rhs_hu", "rhs_hv"), r):
        assert same(a, z[k]), k


@pytest.mark.parametrize("name", ["steps_gauss_64_10", "steps_gauss_64_100", "steps_lake_48_10",
                                  "steps_random_20x24_1", "steps_wetdry_48_3", "steps_wetdry_48_10"])
def test_fixed_dt_steps(name):
    m, z = MANIFEST[name], load(name)
    gr = oracle.grid(m["nx"], m["ny"], m["ng"], m["dx"], m["dy"])
    h, hu, hv = z["h"].copy(), z["hu"].copy(), z["hv"].copy()
    dt0, _ = oracle.stable_dt(gr, h, hu, hv, m["cfl"])
    assert same([dt0], z["dt_stable"])
    rc, _, _ = oracle.run(gr, h, hu, hv, z["B"], m["steps"], dt=float(z["dt"][0]))
    assert rc == m["rc"]
    for k, a in zip(("out_h", "out_hu", "out_hv"), (h, hu, hv)):
        assert same(a, z[k]), k


def test_adaptive_steps():
    name = "adaptive_gauss_48_20"
    m, z = MANIFEST[name], load(name)
    gr = oracle.grid(m["nx"], m["ny"], m["ng"], m["dx"], m["dy"])
    h, hu, hv = z["h"].copy(), z["hu"].copy(), z["hv"].copy()
    rc, _, hist = oracle.run(gr, h, hu, hv, z["B"], m["steps"], cfl=m["cfl"])
    assert rc == 0
    assert same(hist, z["dt_hist"])
    for k, a in zip(("out_h", "out_hu", "out_hv"), (h, hu, hv)):
        assert same(a, z[k]), k


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.slow
def test_gauss_256_1000_hash():
    """C1 geometry (256^2, 1000 steps) at the reference g: final fields by sha256."""
    m = MANIFEST["steps_gauss_256_1000"]
    c = W.gauss_pulse(256, g=oracle.G_REF)
    gr = oracle.grid(256, 256, 2, c["dx"], c["dy"])
    h, hu, hv, B = c["h"].copy(), c["hu"].copy(), c["hv"].copy(), c["B"].copy()
    oracle.apply_bc(gr, B)
    oracle.apply_bc(gr, h, hu, hv)
    for k, v in {"h": h, "hu": hu, "hv": hv, "B": B}.items():
        if sha(v) != m["sha256_in"][k]:
            pytest.skip(f"generated input {k} differs from the fixture's (libm exp differs)")
    dt, _ = oracle.stable_dt(gr, h, hu, hv, 0.4)
    assert dt == m["dt_stable"]
    rc, _, _ = oracle.run(gr, h, hu, hv, B, 1000, dt=dt)
    assert rc == 0
    for k, a in zip(("h", "hu", "hv"), (h, hu, hv)):
        assert sha(a) == m["sha256_out"][k], k
