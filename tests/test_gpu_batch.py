"""Batched ensemble: every instance bitwise equal to the plain single-grid path (and, for a
few, to the oracle), for fixed and adaptive dt, any launch shape; per-instance flags."""
import numpy as np
import pytest

import oracle
from swe_hip import workloads as W

pytestmark = pytest.mark.gpu

SIDES = [5, 64, 61, 130, 7, 200, 33, 2, 97]


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def same(a, b):
    return np.array_equal(bits(a), bits(b))


def members(sides=SIDES, bc=0):
    cases = []
    for i, n in enumerate(sides):
        c = W.ensemble_member(i, n)
        c["bc"] = bc
        cases.append(c)
    return cases


def make_batch(hip, cases):
    b = hip.HipBatch(cases)
    for i, c in enumerate(cases):
        b.upload(i, c["h"], c["hu"], c["hv"], c["B"])
    return b


def single_run(hip, c, nsteps, dt=None, cfl=None):
    with hip.HipSolver.from_case(c) as s:
        s.upload(c["h"], c["hu"], c["hv"], c["B"])
        hist = None
        if dt is not None:
            s.step(dt, nsteps)
        else:
            hist = s.step_adaptive(cfl, nsteps)
        return s.download(), hist


@pytest.mark.parametrize("bc", [0, 1])
def test_batch_fixed_dt_matches_single(hip, bc):
    cases = members(bc=bc)
    with make_batch(hip, cases) as b:
        dts = b.stable_dt(0.4)
        b.step(13)
        got = [b.download(i) for i in range(len(cases))]
    for i, c in enumerate(cases):
        want, _ = single_run(hip, c, 13, dt=float(dts[i]))
        for comp, g_, w_ in zip(("h", "hu", "hv"), got[i], want):
            assert same(g_, w_), f"instance {i} (n={c['nx']}) {comp}"


def test_batch_stable_dt_matches_oracle(hip):
    cases = members()
    with make_batch(hip, cases) as b:
        dts = b.stable_dt(0.4)
    for i, c in enumerate(cases):
        gr = oracle.grid(c["nx"], c["ny"], 2, c["dx"], c["dy"], g=c["g"])
        want, _ = oracle.stable_dt(gr, c["h"], c["hu"], c["hv"], 0.4)
        assert dts[i] == want


def test_batch_adaptive_matches_single_and_oracle(hip):
    cases = members()
    with make_batch(hip, cases) as b:
        hist = b.step_adaptive(0.4, 17)
        got = [b.download(i) for i in range(len(cases))]
    for i, c in enumerate(cases):
        want, whist = single_run(hip, c, 17, cfl=0.4)
        assert same(hist[:, i], whist), f"dt history of instance {i}"
        for comp, g_, w_ in zip(("h", "hu", "hv"), got[i], want):
            assert same(g_, w_), f"instance {i} {comp}"
    for i in (1, 5):
        c = cases[i]
        gr = oracle.grid(c["nx"], c["ny"], 2, c["dx"], c["dy"], g=c["g"])
        B = c["B"].copy()
        oracle.apply_bc(gr, B)
        h, hu, hv = (c[k].copy() for k in ("h", "hu", "hv"))
        rc, _, ohist = oracle.run(gr, h, hu, hv, B, 17, cfl=0.4)
        assert rc == 0 and same(ohist, hist[:, i])
        for g_, w_ in zip(got[i], (h, hu, hv)):
            assert same(g_, w_)


@pytest.mark.parametrize("tile_rows,waves", [(1, 0), (5, 3), (16, 8), (64, 0), (300, 1)])
def test_batch_launch_shape_invariant(hip, tile_rows, waves):
    cases = members(sides=[70, 9, 130, 61])
    ref = None
    for shape in [(32, 0), (tile_rows, waves)]:
        with make_batch(hip, cases) as b:
            b.set_launch(*shape)
            b.step_adaptive(0.4, 6)
            out = [b.download(i) for i in range(len(cases))]
        if ref is None:
            ref = out
        else:
            for i in range(len(cases)):
                for g_, w_ in zip(out[i], ref[i]):
                    assert same(g_, w_), f"shape {shape} instance {i}"


def test_batch_autotune_keeps_state(hip):
    cases = members(sides=[100, 40, 250])
    with make_batch(hip, cases) as b:
        before = [b.download(i) for i in range(len(cases))]
        t, w = b.autotune()
        assert t in (8, 16, 32, 64) and w in (0, 8, 12, 16, 24)
        assert b.get_launch() == (t, w)
        after = [b.download(i) for i in range(len(cases))]
    for x, y in zip(before, after):
        for g_, w_ in zip(x, y):
            assert same(g_, w_)


def test_batch_nonfinite_flags_per_instance(hip):
    """A non-finite instance stops at its failing stage as a single grid does (here stage 1
    of step 1: its arrays stay as uploaded) while the others take all their steps, equal to
    their single-grid runs."""
    cases = members(sides=[20, 30, 40])
    cases[1]["hv"][10, 10] = np.nan
    with make_batch(hip, cases) as b:
        b.set_dt(np.array([1e-4, 1e-4, 1e-4]))
        with pytest.raises(hip.SweError) as e:
            b.step(2)
        assert e.value.code == hip.SWE_ERR_NONFINITE
        assert "instance 1 at stage 1 of step 1 " in str(e.value)
        assert list(b.nonfinite()) == [0, 1, 0]
        got = [b.download(i) for i in range(len(cases))]
    I = (slice(2, -2), slice(2, -2))
    for comp, g_ in zip(("h", "hu", "hv"), got[1]):
        assert same(g_[I], cases[1][comp][I]), f"stopped instance {comp}"
    for i in (0, 2):
        want, _ = single_run(hip, cases[i], 2, dt=1e-4)
        for comp, g_, w_ in zip(("h", "hu", "hv"), got[i], want):
            assert same(g_, w_), f"instance {i} {comp}"


def test_batch_adaptive_stop_per_instance(hip):
    """Adaptive dt: the instance that goes non-finite (a rough wet/dry state at CFL 3.0,
    stage 2 of step 15 in the oracle) stops there with the oracle's dt history up to that
    step and quiet NaN after it, and the oracle's arrays; the other instance's history and
    arrays equal its single-grid run."""
    bad = W.random_state(16, 19, seed=2, wet_dry=True)
    good = W.random_state(16, 19, seed=2, wet_dry=False)
    cases = [good, bad]
    with make_batch(hip, cases) as b:
        with pytest.raises(hip.SweError) as e:
            b.step_adaptive(3.0, 30)
        assert e.value.code == hip.SWE_ERR_NONFINITE
        assert "instance 1 at stage 2 of step 15 " in str(e.value)
        assert list(b.nonfinite()) == [0, 2]
        hist = b.last_dt_hist
        got = [b.download(i) for i in range(2)]
    gr = oracle.grid(16, 19, 2, bad["dx"], bad["dy"], g=bad["g"], bc=bad["bc"])
    B = bad["B"].copy()
    oracle.apply_bc(gr, B)
    U = [bad[k].copy() for k in ("h", "hu", "hv")]
    rc, done, ohist, st = oracle.run_status(gr, *U, B, 30, cfl=3.0)
    assert (rc, done + 1, st) == (1, 15, 2)
    assert same(hist[:15, 1], ohist[:15]) and np.isnan(hist[15:, 1]).all()
    I = (slice(2, -2), slice(2, -2))
    for comp, g_, w_ in zip(("h", "hu", "hv"), got[1], U):
        assert same(g_[I], w_[I]), f"stopped instance {comp}"
    want, whist = single_run(hip, good, 30, cfl=3.0)
    assert same(hist[:, 0], whist)
    for comp, g_, w_ in zip(("h", "hu", "hv"), got[0], want):
        assert same(g_, w_), f"running instance {comp}"


def test_batch_rejects_mixed_parameters(hip):
    a, c = W.ensemble_member(0, 10), W.ensemble_member(1, 12)
    c["g"] = 9.81
    with pytest.raises(hip.SweError) as e:
        hip.HipBatch([a, c])
    assert e.value.code == hip.SWE_ERR_INVALID


@pytest.mark.slow
def test_ensemble_512_full_size_bitwise(hip):
    """Config 4 at its real size: 512 instances (sides log-uniform in [64, 1024], seed 4,
    96.6 M cells), 500 steps with per-instance adaptive dt in one batch (autotuned launch
    shape) against each instance advanced alone: every instance bitwise equal, and the
    per-instance dt sequences with it."""
    sides = W.ensemble_sides(512)
    cases = [W.ensemble_member(i, int(n)) for i, n in enumerate(sides)]
    b = make_batch(hip, cases)
    try:
        b.autotune()   # times stages 1-2 only: the state is untouched
        hist = b.step_adaptive(0.4, 500)
        for i, c in enumerate(cases):
            with hip.HipSolver.from_case(c) as s:
                s.upload(c["h"], c["hu"], c["hv"], c["B"])
                want_hist = s.step_adaptive(0.4, 500)
                want = s.download()
            got = b.download(i)
            assert same(hist[:, i], want_hist), (i, int(sides[i]), "dt")
            assert all(same(x, y) for x, y in zip(got, want)), (i, int(sides[i]))
    finally:
        b.close()
