"""Fused steps (swe_hip_set_step_fusion: the three SSP-RK3 stages of a step in one launch,
temporal blocking): bitwise equal to the three-launch path and to the oracle, for fixed and
adaptive dt, both wall types, odd shapes and tile heights, across calls (the two state
buffers alternate), and with the same first-bad-stage stop.  Runs on an MI355X."""
import numpy as np
import pytest

import oracle
from swe_hip import workloads as W
from test_oracle_invariants import STOP_CASES, stop_case

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def same(a, b):
    return np.array_equal(bits(a), bits(b))


def solver(hip, c, fused_rows=0):
    s = hip.HipSolver.from_case(c)
    s.upload(c["h"], c["hu"], c["hv"], c["B"])
    s.step_fusion = fused_rows
    return s


SHAPES = [(2, 2), (3, 53), (17, 5), (40, 104), (61, 105), (130, 52), (9, 130)]


def run_steps(hip, c, rows, dt, n):
    """(status of the call: (steps, stage), arrays) of n steps in one call."""
    with solver(hip, c, rows) as s:
        try:
            s.step(dt, n)
        except hip.SweError as e:
            assert e.code == hip.SWE_ERR_NONFINITE
        return s.last_status(), s.download()


def same_or_both_nan(a, b):
    return not ((bits(a) != bits(b)) & ~(np.isnan(a) & np.isnan(b))).any()


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("bc", [0, 1])
def test_fused_equals_three_launches_and_oracle(hip, shape, bc):
    """5 steps of a rough random state (wet/dry on odd-sum shapes): the three-launch path
    matches the oracle (the final arrays, or the step, stage and arrays where the run goes
    non-finite), and the fused path with 4-, 5-, 16- and 64-row tiles matches it bitwise."""
    nx, ny = shape
    c = W.random_state(nx, ny, seed=31 + nx + 7 * ny, wet_dry=bool((nx + ny) % 2), bc=bc)
    gr = oracle.grid(nx, ny, 2, c["dx"], c["dy"], g=c["g"], bc=bc)
    B = c["B"].copy()
    oracle.apply_bc(gr, B)
    U = [c[k].copy() for k in ("h", "hu", "hv")]
    dt, _ = oracle.stable_dt(gr, *U, 0.3)
    rc, done, _, stage = oracle.run_status(gr, *U, B, 5, dt=dt)
    want_status, want = run_steps(hip, c, 0, dt, 5)
    I = (slice(2, -2), slice(2, -2))
    if rc == 0:
        assert want_status == (5, 0)
        for comp, g_, w_ in zip(("h", "hu", "hv"), want, U):
            assert same(g_, w_), ("three launches vs oracle", comp)
    else:
        assert want_status == (done + 1, stage)
        for comp, g_, w_ in zip(("h", "hu", "hv"), want, U):
            assert same_or_both_nan(g_[I], w_[I]), ("three launches vs oracle", comp)
    for rows in (4, 5, 16, 64):
        status, got = run_steps(hip, c, rows, dt, 5)
        assert status == want_status, (shape, bc, rows)
        for comp, g_, w_ in zip(("h", "hu", "hv"), got, want):
            assert same_or_both_nan(g_, w_), (shape, bc, rows, comp)


def test_fused_adaptive_history_and_state(hip):
    c = W.gauss_pulse(96, g=hip.G_REF)
    with solver(hip, c) as s:
        want_hist = s.step_adaptive(0.45, 12)
        want = s.download()
    with solver(hip, c, 16) as s:
        hist = s.step_adaptive(0.45, 12)
        got = s.download()
    assert same(hist, want_hist)
    for g_, w_ in zip(got, want):
        assert same(g_, w_)


@pytest.mark.parametrize("bc", [0, 1])
def test_fused_fast_precision_and_adaptive_transmissive(hip, bc):
    """The FAST precision mode (not bitwise to the reference, but the same arithmetic in
    both launch modes) and adaptive dt with either wall type: the fused step equals the
    three launches bitwise."""
    c = W.random_state(77, 131, seed=8, bc=bc)
    runs = []
    for rows in (0, 12):
        with solver(hip, c, rows) as s:
            s.precision = hip.PRECISION_FAST
            hist = s.step_adaptive(0.35, 7)
            runs.append((hist, s.download()))
    (h0, a0), (h1, a1) = runs
    assert same(h0, h1)
    for g_, w_ in zip(a1, a0):
        assert same(g_, w_)


def test_fused_across_calls_and_modes(hip):
    """Odd and even step counts (the state buffers swap every step), then back to three
    launches (a fresh graph on the current buffers), then fused again."""
    c = W.random_state(70, 90, seed=5)
    with solver(hip, c) as s:
        dt, _ = s.stable_dt(0.3)
        s.step(dt, 12)
        want = s.download()
    with solver(hip, c, 8) as s:
        s.step(dt, 3)
        s.step(dt, 2)
        s.step_fusion = 0
        s.step(dt, 4)
        s.step_fusion = 32
        s.step(dt, 1)
        s.step(dt, 2)
        got = s.download()
    for g_, w_ in zip(got, want):
        assert same(g_, w_)


def test_fused_mixed_with_other_entry_points(hip):
    """Zero-step calls, the allocation-free adaptive step and the right-hand side between
    fused calls all act on the current state buffer."""
    torch = pytest.importorskip("torch")
    c = W.gauss_pulse(80, g=hip.G_REF)
    with solver(hip, c) as s:
        s.step(2e-4, 3)
        h1 = s.step_adaptive(0.4, 4)
        r1 = s.rhs()
        s.step(2e-4, 2)
        want = s.download()
    with solver(hip, c, 16) as s:
        s.step(2e-4, 0)
        s.step(2e-4, 3)
        hist = torch.zeros(4, dtype=torch.float64, device="cuda")
        s.step_adaptive_noalloc(0.4, 4, hist)
        r2 = s.rhs()
        s.step(2e-4, 2)
        got = s.download()
    assert same(hist.cpu().numpy(), h1)
    for a, b in zip(r2, r1):
        assert same(a, b)
    for g_, w_ in zip(got, want):
        assert same(g_, w_)


@pytest.mark.parametrize("name", sorted(STOP_CASES))
def test_fused_first_bad_stage_stop(hip, name):
    """The fused launch that finds a non-finite stage output runs to its end, every later
    launch of the call does nothing, and the call reports the stage and step the oracle
    stops at and leaves the same arrays (U^n, or the stage-3 output) -- with an odd and an
    even number of steps left after the failure."""
    c, gr, B, U, dt, stage, step = stop_case(name)
    rc, done, _, st = oracle.run_status(gr, *U, B, 20, dt=dt)
    assert (rc, done + 1, st) == (1, step, stage)
    I = (slice(2, -2), slice(2, -2))
    for nsteps in (step + 2, step + 3):
        with solver(hip, c, 8) as s:
            with pytest.raises(hip.SweError) as e:
                s.step(dt, nsteps)
            assert e.value.code == hip.SWE_ERR_NONFINITE
            assert s.last_status() == (step, stage)
            got = s.download()
        for comp, g_, w_ in zip(("h", "hu", "hv"), got, U):
            neq = (bits(g_[I]) != bits(w_[I])) & ~(np.isnan(g_[I]) & np.isnan(w_[I]))
            assert not neq.any(), (name, nsteps, comp)


def test_fused_rejects_ghost_width_3(hip):
    c = W.random_state(20, 20, seed=1, ng=3)
    with hip.HipSolver.from_case(c) as s:
        with pytest.raises(hip.SweError) as e:
            s.step_fusion = 16
        assert e.value.code == hip.SWE_ERR_INVALID


@pytest.mark.slow
def test_fused_dam_break_8192(hip):
    """Config 3's grid: 3 fused steps (64-row tiles) bitwise equal to 3 three-launch steps."""
    c = W.dam_break(8192)
    with solver(hip, c) as s:
        dt, _ = s.stable_dt(0.45)
        s.step(dt, 3)
        want = s.download()
    with solver(hip, c, 64) as s:
        s.step(dt, 3)
        got = s.download()
    for g_, w_ in zip(got, want):
        assert same(g_, w_)
