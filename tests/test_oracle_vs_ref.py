"""The oracle against the reference's own kernels compiled here (oracle/_ref), on fresh
random inputs.  Skipped where the reference harness was not built (e.g. a tree without
the reference).  Bar: bitwise."""
import numpy as np
import pytest

import oracle
from swe_hip import workloads as W

pytestmark = pytest.mark.skipif(not oracle.ref_available(),
                                reason="oracle/_ref not built (reference tree absent)")


def same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64),
                          np.ascontiguousarray(b).view(np.uint64))


def test_ref_constant():
    assert oracle.ref().ref_g() == oracle.G_REF


def test_minmod_random_and_special():
    R = oracle.ref()
    vals = np.concatenate([W.uniform(1, 500, 0, -3, 3), [0.0, -0.0, 1e-170, -1e-170, 1e308,
                                                         -1e308, np.inf, -np.inf, 5e-324]])
    for a in vals:
        for b in vals[::7]:
            x, y = oracle.minmod(a, b), R.ref_minmod(a, b)
            assert same(np.array([x]), np.array([y])), (a, b)
            assert R.hpc_minmod(a, b) == y or (np.isnan(y) and np.isnan(R.hpc_minmod(a, b)))


@pytest.mark.parametrize("shape,wd,seed,ng", [((9, 31), False, 1, 2), ((33, 20), True, 2, 2),
                                              ((64, 65), True, 3, 2), ((2, 3), False, 4, 2),
                                              ((21, 14), True, 5, 3)])
def test_faces_rhs_random(shape, wd, seed, ng):
    c = W.random_state(*shape, seed=seed, wet_dry=wd, ng=ng)
    gr = oracle.grid(*shape, ng, c["dx"], c["dy"])
    B = c["B"].copy()
    oracle.apply_bc(gr, B)
    h, hu, hv = c["h"].copy(), c["hu"].copy(), c["hv"].copy()
    oracle.apply_bc(gr, h, hu, hv)
    fo = oracle.faces(gr, h, hu, hv, B)
    fr, rr = oracle.ref_faces_and_rhs(gr, h, hu, hv, B)
    fh, rh = oracle.ref_faces_and_rhs(gr, h, hu, hv, B, hpc=True)
    ro = oracle.rhs(gr, h.copy(), hu.copy(), hv.copy(), B)
    for a, b, c_ in zip(fo, fr, fh):
        assert same(a, b) and same(b, c_)
    for a, b, c_ in zip(ro, rr, rh):
        assert same(a, b) and same(b, c_)


@pytest.mark.parametrize("ng", [2, 3])
def test_bc_fill_matches_reference(ng):
    c = W.random_state(13, 17, seed=9, ng=ng)
    gr = oracle.grid(13, 17, ng, c["dx"], c["dy"])
    a = [c[k].copy() for k in ("h", "hu", "hv")]
    b = [c[k].copy() for k in ("h", "hu", "hv")]
    oracle.apply_bc(gr, *a)
    oracle.ref().ref_reflect(13, 17, ng, gr.dx, gr.dy, *(oracle._p(x) for x in b))
    for x, y in zip(a, b):
        assert same(x, y)
    s1, s2 = c["B"].copy(), c["B"].copy()
    oracle.apply_bc(gr, s1)
    oracle.ref().ref_reflect(13, 17, ng, gr.dx, gr.dy, oracle._p(s2), None, None)
    assert same(s1, s2)


@pytest.mark.parametrize("mode", ["fixed", "adaptive"])
@pytest.mark.parametrize("case", ["gauss_40", "member_70", "lake_wet_36", "wetdry_36",
                                  "random_ng3", "stop_stage2", "stop_stage3"])
def test_runs_match(mode, case):
    c = {"gauss_40": lambda: W.gauss_pulse(40, g=oracle.G_REF),
         "member_70": lambda: W.ensemble_member(2, 70),
         "lake_wet_36": lambda: W.lake_wet_dry(36, seed=3, eta0=2.0),
         "wetdry_36": lambda: W.lake_wet_dry(36, seed=3),
         "random_ng3": lambda: W.random_state(19, 26, seed=6, ng=3),
         # the reference throws in stage 2 / stage 3 (fixed dt, CFL 0.4 / 1.5): the arrays
         # it leaves
         "stop_stage2": lambda: W.random_state(16, 19, seed=1, wet_dry=True),
         "stop_stage3": lambda: W.random_state(16, 19, seed=0)}[case]()
    cfl = 1.5 if case == "stop_stage3" else 0.4
    gr = oracle.grid(c["nx"], c["ny"], c["ng"], c["dx"], c["dy"])
    B = c["B"].copy()
    oracle.apply_bc(gr, B)
    a = [c[k].copy() for k in ("h", "hu", "hv")]
    b = [c[k].copy() for k in ("h", "hu", "hv")]
    if mode == "fixed":
        dt, _ = oracle.stable_dt(gr, *a, cfl)
        rc1, done, _, stage = oracle.run_status(gr, *a, B, 12, dt=dt)
        rc2, _ = oracle.ref_run(gr, *b, B, 12, dt=dt)
        if case.startswith("stop_"):
            assert (rc1, stage) == (1, int(case[-1]))
    else:
        rc1, done, h1, _ = oracle.run_status(gr, *a, B, 12, cfl=cfl)
        rc2, h2 = oracle.ref_run(gr, *b, B, 12, cfl=cfl)
        n = 12 if rc1 == 0 else done + 1  # the dts up to the throw
        assert same(h1[:n], h2[:n])
    assert rc1 == rc2
    for x, y in zip(a, b):
        assert same(x, y)


def test_positivity_and_nonfinite_match():
    c = W.random_state(11, 12, seed=5, wet_dry=True)
    gr = oracle.grid(11, 12, 2, c["dx"], c["dy"])
    a = [c[k].copy() for k in ("h", "hu", "hv")]
    a[0][3, 3] = -1.0
    b = [x.copy() for x in a]
    assert oracle.enforce_positivity(gr, *a) == 0
    assert oracle.ref().ref_enforce_positivity(11, 12, 2, gr.dx, gr.dy,
                                               *(oracle._p(x) for x in b)) == 0
    for x, y in zip(a, b):
        assert same(x, y)
    a[1][4, 4] = np.nan
    b[1][4, 4] = np.nan
    assert oracle.enforce_positivity(gr, *a) == 1
    assert oracle.ref().ref_enforce_positivity(11, 12, 2, gr.dx, gr.dy,
                                               *(oracle._p(x) for x in b)) == 1
