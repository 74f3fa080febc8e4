"""GPU parity: the HIP fused-stage path against the oracle and the reference's golden vectors.

Bar: bitwise equality (every float64 bit, ghosts included) -- the HIP kernels keep the
reference's operation order and are compiled without FP contraction.  Runs on an MI355X
through the C ABI (swe_hip -> libswe_hip.so)."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
from swe_hip import workloads as W
from test_oracle_invariants import STOP_CASES, stop_case

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def assert_bitwise(got, want, what="", nan_any=False):
    """Every float64 bit equal.  nan_any: a NaN matches any NaN -- IEEE 754 leaves the sign and
    payload of the NaN an invalid operation produces to the hardware (x86 gives -NaN, the
    GPU +NaN), so states that have gone non-finite compare NaN positions only."""
    got, want = np.asarray(got), np.asarray(want)
    assert got.shape == want.shape, what
    neq = bits(got) != bits(want)
    if nan_any:
        neq &= ~(np.isnan(got) & np.isnan(want))
    if neq.any():
        idx = np.argwhere(neq)
        d = np.abs(got - want)
        pytest.fail(f"{what}: {neq.sum()} of {neq.size} values differ, first at {tuple(idx[0])}: "
                    f"got {got[tuple(idx[0])]!r} want {want[tuple(idx[0])]!r}, max |diff| "
                    f"{np.nanmax(d):.3e}")


def oracle_grid(c, bc=None):
    return oracle.grid(c["nx"], c["ny"], c["ng"], c["dx"], c["dy"], g=c["g"],
                       bc=c["bc"] if bc is None else bc)


def gpu_solver(hip, c):
    s = hip.HipSolver.from_case(c)
    s.upload(c["h"], c["hu"], c["hv"], c["B"])
    return s


def oracle_state(c):
    gr = oracle_grid(c)
    B = c["B"].copy()
    oracle.apply_bc(gr, B)
    h, hu, hv = (c[k].copy() for k in ("h", "hu", "hv"))
    return gr, h, hu, hv, B


RHS_CASES = [
    ("random_17x13", lambda: W.random_state(17, 13, seed=11)),
    ("wetdry_32x32", lambda: W.random_state(32, 32, seed=12, wet_dry=True)),
    ("random_61x130", lambda: W.random_state(61, 130, seed=14)),
    ("random_130x61_wd", lambda: W.random_state(130, 61, seed=15, wet_dry=True)),
    ("random_2x2", lambda: W.random_state(2, 2, seed=16)),
    ("random_5x121_transm", lambda: W.random_state(5, 121, seed=17, bc=1)),
    ("lake_64_transm", lambda: W.lake_wet_dry(64, seed=2)),
]


@pytest.mark.parametrize("name,make", RHS_CASES, ids=[c[0] for c in RHS_CASES])
def test_rhs_bitwise(hip, name, make):
    c = make()
    gr, h, hu, hv, B = oracle_state(c)
    want = oracle.rhs(gr, h, hu, hv, B)
    with gpu_solver(hip, c) as s:
        got = s.rhs()
    for comp, g_, w_ in zip(("h", "hu", "hv"), got, want):
        assert_bitwise(g_, w_, f"{name} rhs_{comp}")


@pytest.mark.parametrize("fixture", ["rhs_random_17x13", "rhs_wetdry_32x32", "rhs_lake_24"])
def test_rhs_golden(hip, fixture):
    """rhs of the reference's compiled kernels (tests/golden), straight against the GPU."""
    m = json.load(open(os.path.join(GOLDEN, "manifest.json")))[fixture]
    z = np.load(os.path.join(GOLDEN, fixture + ".npz"))
    with hip.HipSolver(m["nx"], m["ny"], m["dx"], m["dy"], hip.G_REF, 0, m["ng"]) as s:
        s.upload(z["h"], z["hu"], z["hv"], z["B"])
        got = s.rhs()
    for comp, g_ in zip(("h", "hu", "hv"), got):
        assert_bitwise(g_, z["rhs_" + comp], f"{fixture} rhs_{comp}")


@pytest.mark.parametrize("fixture", ["steps_gauss_64_10", "steps_gauss_64_100", "steps_lake_48_10",
                                     "steps_random_20x24_1", "steps_wetdry_48_3"])
def test_steps_golden(hip, fixture):
    m = json.load(open(os.path.join(GOLDEN, "manifest.json")))[fixture]
    z = np.load(os.path.join(GOLDEN, fixture + ".npz"))
    assert m["rc"] == 0
    with hip.HipSolver(m["nx"], m["ny"], m["dx"], m["dy"], hip.G_REF, 0, m["ng"]) as s:
        s.upload(z["h"], z["hu"], z["hv"], z["B"])
        dt, _ = s.stable_dt(m["cfl"])
        assert bits(np.array([dt]))[0] == bits(z["dt_stable"])[0]
        s.step(float(z["dt"][0]), m["steps"])
        got = s.download()
    for comp, g_ in zip(("h", "hu", "hv"), got):
        assert_bitwise(g_, z["out_" + comp], f"{fixture} {comp}")


def test_adaptive_golden(hip):
    fixture = "adaptive_gauss_48_20"
    m = json.load(open(os.path.join(GOLDEN, "manifest.json")))[fixture]
    z = np.load(os.path.join(GOLDEN, fixture + ".npz"))
    with hip.HipSolver(m["nx"], m["ny"], m["dx"], m["dy"], hip.G_REF, 0, m["ng"]) as s:
        s.upload(z["h"], z["hu"], z["hv"], z["B"])
        hist = s.step_adaptive(m["cfl"], m["steps"])
        got = s.download()
    assert_bitwise(hist, z["dt_hist"], "dt history")
    for comp, g_ in zip(("h", "hu", "hv"), got):
        assert_bitwise(g_, z["out_" + comp], f"{fixture} {comp}")


def test_gauss_256_1000_reference_hash(hip):
    """C1 geometry at g = 9.80665, 1000 steps: sha256 of the reference's final fields."""
    m = json.load(open(os.path.join(GOLDEN, "manifest.json")))["steps_gauss_256_1000"]
    c = W.gauss_pulse(256, g=hip.G_REF)
    gr, h, hu, hv, B = oracle_state(c)
    oracle.apply_bc(gr, h, hu, hv)
    got_in = {"h": h, "hu": hu, "hv": hv, "B": B}
    for k, v in got_in.items():
        if hashlib.sha256(v.tobytes()).hexdigest() != m["sha256_in"][k]:
            pytest.skip(f"generated input {k} differs from the fixture's (libm exp differs)")
    with gpu_solver(hip, c) as s:
        dt, _ = s.stable_dt(0.4)
        assert dt == m["dt_stable"]
        s.step(dt, 1000)
        out = s.download()
    for comp, a in zip(("h", "hu", "hv"), out):
        assert hashlib.sha256(a.tobytes()).hexdigest() == m["sha256_out"][comp], comp


def test_gauss_256_g981_vs_oracle(hip):
    """Config 1 as stated (g = 9.81, CFL 0.4 -> fixed dt, 1000 steps), bitwise vs oracle."""
    c = W.gauss_pulse(256, g=9.81)
    gr, h, hu, hv, B = oracle_state(c)
    dt_o, _ = oracle.stable_dt(gr, h, hu, hv, 0.4)
    with gpu_solver(hip, c) as s:
        dt, _ = s.stable_dt(0.4)
        assert dt == dt_o
        s.step(dt, 1000)
        got = s.download()
    rc, _, _ = oracle.run(gr, h, hu, hv, B, 1000, dt=dt_o)
    assert rc == 0
    for comp, g_, w_ in zip(("h", "hu", "hv"), got, (h, hu, hv)):
        assert_bitwise(g_, w_, f"gauss_256 g=9.81 {comp}")


@pytest.mark.parametrize("bc", [0, 1])
def test_adaptive_vs_oracle(hip, bc):
    c = W.ensemble_member(3, 150)
    c["bc"] = bc
    gr, h, hu, hv, B = oracle_state(c)
    with gpu_solver(hip, c) as s:
        hist = s.step_adaptive(0.4, 40)
        got = s.download()
    rc, _, want_hist = oracle.run(gr, h, hu, hv, B, 40, cfl=0.4)
    assert rc == 0
    assert_bitwise(hist, want_hist, "dt history")
    for comp, g_, w_ in zip(("h", "hu", "hv"), got, (h, hu, hv)):
        assert_bitwise(g_, w_, f"adaptive bc={bc} {comp}")


@pytest.mark.parametrize("flat", [False, True], ids=["bathy", "flat_bottom"])
@pytest.mark.parametrize("case", ["tiny_values", "huge_values", "denormals"])
def test_split_guard_fallback(hip, case, flat):
    """Cells outside [2^-240, 2^190] (and not 0) switch their waves to the full IEEE division
    and sqrt sequences; results stay bitwise equal either way (flat: B = +0.0 everywhere,
    through the kernels that do not read it)."""
    c = W.random_state(90, 140, seed=31, bathy=not flat)
    h, hu, hv, B = c["h"], c["hu"], c["hv"], c["B"]
    if case == "tiny_values":
        hu[10:14, 20:70] = 1e-300
        hv[40:42, 100:130] = -3e-250
        if not flat:
            B[60:63, 5:15] = 1e-290
    elif case == "huge_values":
        hv[30:33, 50:60] = 1e60   # above 2^190 (~1.6e57)
        if not flat:
            B[70, 70:90] = 1e58
        else:
            h[70, 70:90] = 1e58
    else:
        hu[20:25, 80:90] = 5e-320
        hv[50:52, 10:40] = -1e-310
    gr, h0, hu0, hv0, B0 = oracle_state(c)
    want = oracle.rhs(gr, h0.copy(), hu0.copy(), hv0.copy(), B0)
    with gpu_solver(hip, c) as s:
        assert s.flat_bottom == flat
        got = s.rhs()
        dt, _ = s.stable_dt(0.2)
        s.step(dt, 2)
        got_state = s.download()
    for comp, g_, w_ in zip(("h", "hu", "hv"), got, want):
        assert_bitwise(g_, w_, f"{case} rhs_{comp}")
    rc, _, _ = oracle.run(gr, h0, hu0, hv0, B0, 2, dt=dt)
    if rc == 0:
        for comp, g_, w_ in zip(("h", "hu", "hv"), got_state, (h0, hu0, hv0)):
            assert_bitwise(g_, w_, f"{case} 2 steps {comp}")


@pytest.mark.parametrize("tile_rows", [1, 3, 16, 64, 1000])
def test_tile_rows_invariant(hip, tile_rows):
    c = W.lake_wet_dry(97, seed=6, eta0=2.0)
    gr, h, hu, hv, B = oracle_state(c)
    dt, _ = oracle.stable_dt(gr, h, hu, hv, 0.4)
    with gpu_solver(hip, c) as s:
        s.tile_rows = tile_rows
        s.step(dt, 7)
        got = s.download()
    oracle.run(gr, h, hu, hv, B, 7, dt=dt)
    for comp, g_, w_ in zip(("h", "hu", "hv"), got, (h, hu, hv)):
        assert_bitwise(g_, w_, f"tile_rows={tile_rows} {comp}")


@pytest.mark.parametrize("rows", [(32, 24, 24), (5, 64, 3), (1000, 1, 17)])
def test_stage_tile_rows_invariant(hip, rows):
    """Each stage launch with its own tile height: fixed and adaptive steps bitwise equal
    to the oracle, and the graph recaptured when the heights change."""
    c = W.lake_wet_dry(97, seed=6, eta0=2.0)
    gr, h, hu, hv, B = oracle_state(c)
    dt, _ = oracle.stable_dt(gr, h, hu, hv, 0.4)
    with gpu_solver(hip, c) as s:
        s.step(dt, 2)
        s.stage_tile_rows = rows
        assert s.stage_tile_rows == tuple(min(r, c["nx"]) for r in rows)
        assert s.tile_rows == min(rows[0], c["nx"])
        s.step(dt, 5)
        got = s.download()
    oracle.run(gr, h, hu, hv, B, 7, dt=dt)
    for comp, g_, w_ in zip(("h", "hu", "hv"), got, (h, hu, hv)):
        assert_bitwise(g_, w_, f"stage rows {rows} {comp}")
    hists = []
    for r in (rows, (8, 8, 8)):
        with gpu_solver(hip, c) as s:
            s.stage_tile_rows = r
            hists.append((s.step_adaptive(0.4, 6), s.download()))
    assert np.array_equal(hists[0][0], hists[1][0])
    for a, b in zip(hists[0][1], hists[1][1]):
        assert_bitwise(a, b, f"adaptive, stage rows {rows} vs 8")


FLAT_CASES = [
    ("random_37x70", lambda: W.random_state(37, 70, seed=31, bathy=False)),
    ("random_70x37_wd", lambda: W.random_state(70, 37, seed=32, bathy=False, wet_dry=True)),
    ("random_41x129_transm", lambda: W.random_state(41, 129, seed=33, bathy=False, bc=1)),
    ("random_33x33_wd_transm", lambda: W.random_state(33, 33, seed=34, bathy=False,
                                                      wet_dry=True, bc=1)),
    ("dam_break_97", lambda: W.dam_break(97)),
    ("gauss_pulse_64", lambda: W.gauss_pulse(64)),
]


def _gpu_rc(hip, fn):
    """0 when fn() returns, 1 when it raises the non-finite stop (SWE_ERR_NONFINITE)."""
    try:
        fn()
        return 0
    except hip.SweError as e:
        assert e.code == 3, e
        return 1


@pytest.mark.parametrize("name,make", FLAT_CASES, ids=[c[0] for c in FLAT_CASES])
def test_flat_bottom_kernels(hip, name, make):
    """A bottom of +0.0 everywhere: the stage kernels that do not read B give the oracle's
    bits (right-hand side, fixed and adaptive steps, and the arrays a non-finite stop
    leaves) and those of the kernels that do."""
    c = make()
    gr, h, hu, hv, B = oracle_state(c)
    want_rhs = oracle.rhs(gr, h, hu, hv, B)
    dt, _ = oracle.stable_dt(gr, h, hu, hv, 0.4)
    ha, hua, hva = h.copy(), hu.copy(), hv.copy()
    rc, _, _ = oracle.run(gr, h, hu, hv, B, 5, dt=dt)
    rca, _, want_hist = oracle.run(gr, ha, hua, hva, B, 4, cfl=0.4)
    outs = []
    for flat in (True, False):
        with gpu_solver(hip, c) as s:
            s.set_flat_bottom(flat)
            assert s.flat_bottom == flat
            rhs = s.rhs()
            assert _gpu_rc(hip, lambda: s.step(dt, 5)) == rc, f"{name} stop"
            got = s.download()
        with gpu_solver(hip, c) as s:
            s.set_flat_bottom(flat)
            assert _gpu_rc(hip, lambda: s.step_adaptive(0.4, 4)) == rca, f"{name} stop"
            hist = s.last_dt_hist
            got_a = s.download()
        for comp, g_, w_ in zip(("h", "hu", "hv"), rhs, want_rhs):
            assert_bitwise(g_, w_, f"{name} flat={flat} rhs {comp}")
        for comp, g_, w_ in zip(("h", "hu", "hv"), got, (h, hu, hv)):
            assert_bitwise(g_, w_, f"{name} flat={flat} 5 steps {comp}", nan_any=True)
        if rca == 0:
            assert np.array_equal(bits(hist), bits(np.asarray(want_hist))), f"{name} dts"
        for comp, g_, w_ in zip(("h", "hu", "hv"), got_a, (ha, hua, hva)):
            assert_bitwise(g_, w_, f"{name} flat={flat} adaptive {comp}", nan_any=True)
        outs.append(got)
    for a, b in zip(*outs):
        assert_bitwise(a, b, f"{name} flat vs general", nan_any=True)


def test_flat_bottom_detection(hip):
    """Only a bottom whose every bit is zero (after the ghost fill) takes the kernels that do
    not read it: -0.0 or one non-zero cell does not; a new upload re-decides, and the step
    graph follows."""
    c = W.random_state(45, 62, seed=35, bathy=False)
    gr, h, hu, hv, B = oracle_state(c)
    dt, _ = oracle.stable_dt(gr, h, hu, hv, 0.4)
    bumpy = c["B"].copy()
    bumpy[c["ng"] + 20, c["ng"] + 31] = 1e-3
    negzero = np.full_like(c["B"], -0.0)
    ghost_only = c["B"].copy()
    ghost_only[0, 5] = 0.25  # a ghost value: the upload's ghost fill overwrites it
    with gpu_solver(hip, c) as s:
        assert s.flat_bottom
        for Bx, flat in ((bumpy, False), (negzero, False), (ghost_only, True), (c["B"], True),
                         (bumpy, False)):
            s.upload(c["h"], c["hu"], c["hv"], Bx)
            assert s.flat_bottom == flat
            s.step(dt, 3)
            got = s.download()
            Bo = Bx.copy()
            oracle.apply_bc(gr, Bo)
            h2, hu2, hv2 = (c[k].copy() for k in ("h", "hu", "hv"))
            oracle.run(gr, h2, hu2, hv2, Bo, 3, dt=dt)
            for comp, g_, w_ in zip(("h", "hu", "hv"), got, (h2, hu2, hv2)):
                assert_bitwise(g_, w_, f"flat={flat} {comp}")


def test_graph_and_direct_launch_agree(hip):
    c = W.ensemble_member(7, 90)
    outs = []
    for graph in (True, False):
        with gpu_solver(hip, c) as s:
            s.set_graph(graph)
            dt, _ = s.stable_dt(0.4)
            s.step(dt, 5)
            s.step(dt, 6)
            s.step(dt, 70)   # two replays of the 32-step graph and 6 direct steps
            s.step(dt, 32)
            outs.append(s.download())
    for a, b in zip(*outs):
        assert_bitwise(a, b, "graph vs direct")
    # adaptive steps: the dt history is written by the captured steps
    outs, hists = [], []
    for graph in (True, False):
        with gpu_solver(hip, c) as s:
            s.set_graph(graph)
            hists.append(np.concatenate([s.step_adaptive(0.4, 40), s.step_adaptive(0.4, 70)]))
            outs.append(s.download())
    assert_bitwise(hists[0], hists[1], "adaptive dt history, graph vs direct")
    for a, b in zip(*outs):
        assert_bitwise(a, b, "adaptive, graph vs direct")


@pytest.mark.parametrize("g", [9.80665, 9.81])
def test_stable_dt_bitwise(hip, g):
    for seed in range(3):
        c = W.random_state(40 + seed, 70, seed=20 + seed, wet_dry=bool(seed & 1), g=g)
        gr, h, hu, hv, B = oracle_state(c)
        want = oracle.stable_dt(gr, h, hu, hv, 0.45)
        with gpu_solver(hip, c) as s:
            got = s.stable_dt(0.45)
        assert got == want


def test_nonfinite_is_reported(hip):
    c = W.gauss_pulse(64, g=hip.G_REF)
    c["hu"][30, 30] = np.nan
    with gpu_solver(hip, c) as s:
        with pytest.raises(hip.SweError) as e:
            s.step(1e-4, 2)
        assert e.value.code == hip.SWE_ERR_NONFINITE
        assert s.last_status() == (1, 1)
        got = s.download()
    for comp, g_ in zip(("h", "hu", "hv"), got):
        assert_bitwise(g_[2:-2, 2:-2], c[comp][2:-2, 2:-2], f"stopped at stage 1 {comp}")


def test_wetdry_divergence_matches_reference(hip):
    """Wet/dry lake (eta 0.5 over B up to 1.6): the reference scheme diverges after a few
    steps; the GPU reproduces the state bitwise up to then and reports the same failure."""
    m = json.load(open(os.path.join(GOLDEN, "manifest.json")))
    z = np.load(os.path.join(GOLDEN, "steps_wetdry_48_3.npz"))
    assert m["steps_wetdry_48_10"]["rc"] == 1
    dx = m["steps_wetdry_48_3"]["dx"]
    with hip.HipSolver(48, 48, dx, dx, hip.G_REF, 0, 2) as s:
        s.upload(z["h"], z["hu"], z["hv"], z["B"])
        s.step(float(z["dt"][0]), 3)
        got = s.download()
        for comp, g_ in zip(("h", "hu", "hv"), got):
            assert_bitwise(g_, z["out_" + comp], f"wetdry 3 steps {comp}")
        with pytest.raises(hip.SweError) as e:
            s.step(float(z["dt"][0]), 7)
        assert e.value.code == hip.SWE_ERR_NONFINITE


INTERIOR = (slice(2, -2), slice(2, -2))


@pytest.mark.parametrize("name", sorted(STOP_CASES))
def test_first_bad_stage_stop(hip, name):
    """A non-finite stage output stops the call there, as the reference's throw ends
    rk3_step (omp_kernels.hpp:322-330): one 20-step call reports the oracle's failing step
    and stage (one case per stage) and leaves the arrays as the reference does -- U^n of the
    failing step after a stage-1 or stage-2 failure, the floored stage-3 output after a
    stage-3 failure (interior bitwise; the download refills the ghosts).  The next call
    starts afresh and fails in its first step: again at that stage when U^n was kept, at
    stage 1 from the non-finite stage-3 output."""
    c, gr, B, U, dt, stage, step = stop_case(name)
    rc, done, _, st = oracle.run_status(gr, *U, B, 20, dt=dt)
    assert (rc, done + 1, st) == (1, step, stage)
    with gpu_solver(hip, c) as s:
        with pytest.raises(hip.SweError) as e:
            s.step(dt, 20)
        assert e.value.code == hip.SWE_ERR_NONFINITE
        assert f"stage {stage} of step {step} " in str(e.value)
        assert s.last_status() == (step, stage)
        got = s.download()
        for comp, g_, w_ in zip(("h", "hu", "hv"), got, U):
            assert_bitwise(g_[INTERIOR], w_[INTERIOR], f"{name} stopped state {comp}",
                           nan_any=True)
        with pytest.raises(hip.SweError):
            s.step(dt, 5)
        assert s.last_status() == (1, stage if stage != 3 else 1)


def test_stopped_state_is_the_references(hip):
    """The reference's own run of the wet/dry lake (48^2, 10 steps, golden fixture written
    from its compiled kernels) throws in stage 1 of step 5; one HIP call of 10 steps stops
    at the same stage and step with the reference's final arrays (interior bitwise)."""
    m = json.load(open(os.path.join(GOLDEN, "manifest.json")))["steps_wetdry_48_10"]
    z = np.load(os.path.join(GOLDEN, "steps_wetdry_48_10.npz"))
    assert m["rc"] == 1
    with hip.HipSolver(48, 48, m["dx"], m["dy"], hip.G_REF, 0, 2) as s:
        s.upload(z["h"], z["hu"], z["hv"], z["B"])
        with pytest.raises(hip.SweError) as e:
            s.step(float(z["dt"][0]), 10)
        assert e.value.code == hip.SWE_ERR_NONFINITE
        assert s.last_status() == (5, 1)
        got = s.download()
    for comp, g_ in zip(("h", "hu", "hv"), got):
        assert_bitwise(g_[INTERIOR], z["out_" + comp][INTERIOR], f"wetdry stop {comp}",
                       nan_any=True)


def test_first_bad_stage_stop_adaptive(hip):
    """Adaptive dt (CFL 3.0 on a rough wet/dry state): the run goes non-finite in stage 2 of
    step 15; the dt history up to that step equals the oracle's bitwise, later entries are
    quiet NaN (no dt is computed from a stopped state), and the state is U^n of step 15."""
    c = W.random_state(16, 19, seed=2, wet_dry=True)
    gr, h, hu, hv, B = oracle_state(c)
    rc, done, hist, st = oracle.run_status(gr, h, hu, hv, B, 30, cfl=3.0)
    assert (rc, done + 1, st) == (1, 15, 2)
    with gpu_solver(hip, c) as s:
        with pytest.raises(hip.SweError) as e:
            s.step_adaptive(3.0, 30)
        assert e.value.code == hip.SWE_ERR_NONFINITE
        assert s.last_status() == (15, 2)
        got_hist = s.last_dt_hist
        got = s.download()
    assert_bitwise(got_hist[:15], hist[:15], "dt history up to the stop")
    assert np.isnan(got_hist[15:]).all()
    for comp, g_, w_ in zip(("h", "hu", "hv"), got, (h, hu, hv)):
        assert_bitwise(g_[INTERIOR], w_[INTERIOR], f"adaptive stop {comp}", nan_any=True)


def test_lake_at_rest_wet_roundoff(hip):
    """Fully wet lake at rest over 32 Gaussian bumps, transmissive walls: stays at rest to
    round-off and matches the oracle bitwise."""
    c = W.lake_wet_dry(256, seed=2, bump=False, eta0=2.0)
    gr, h, hu, hv, B = oracle_state(c)
    dt, _ = oracle.stable_dt(gr, h, hu, hv, 0.4)
    with gpu_solver(hip, c) as s:
        s.step(dt, 200)
        got = s.download()
    oracle.run(gr, h, hu, hv, B, 200, dt=dt)
    for comp, g_, w_ in zip(("h", "hu", "hv"), got, (h, hu, hv)):
        assert_bitwise(g_, w_, f"lake {comp}")
    I = (slice(2, 258), slice(2, 258))
    eta = got[0][I] + B[I]
    assert np.abs(eta - 2.0).max() < 1e-12
    assert np.abs(got[1][I]).max() < 1e-11 and np.abs(got[2][I]).max() < 1e-11


@pytest.mark.slow
@pytest.mark.parametrize("bump", [False, True])
def test_lake_4096_full_size_properties(hip, bump):
    """Config 2 at its real size: 4096^2 on [0,10]^2, 32 seeded Gaussian bumps, transmissive
    walls, CFL 0.4, 2000 steps -- fully wet (eta 2.0; at eta 0.5 the reference scheme itself
    diverges, test_wetdry_divergence_matches_reference).  Too large for the oracle, so
    size-independent properties: dt equals the oracle's CFL expression bitwise; without the
    bump the lake stays at rest to round-off; with the 0.05 bump the state stays finite and
    positive, the wave stays below its initial amplitude and does not reach the walls in
    2000 steps (0.44 s), so mass is conserved to round-off in both runs."""
    n = 4096
    c = W.lake_wet_dry(n, seed=2, bump=bump, eta0=2.0)
    gr = oracle_grid(c)
    B = c["B"].copy()
    oracle.apply_bc(gr, B)
    dt_o, _ = oracle.stable_dt(gr, c["h"].copy(), c["hu"].copy(), c["hv"].copy(), 0.4)
    I = (slice(2, n + 2), slice(2, n + 2))
    mass0 = c["h"][I].sum()
    with gpu_solver(hip, c) as s:
        dt, _ = s.stable_dt(0.4)
        assert dt == dt_o
        s.step(dt, 2000)
        h, hu, hv = s.download()
    del c
    assert np.isfinite(h).all() and np.isfinite(hu).all() and np.isfinite(hv).all()
    eta = h[I] + B[I]
    if not bump:
        assert np.abs(eta - 2.0).max() < 1e-11
        assert np.abs(hu[I]).max() < 1e-11 and np.abs(hv[I]).max() < 1e-11
    else:
        assert h[I].min() > 0.3
        assert np.abs(eta - 2.0).max() < 0.05
    rim = np.ones((n, n), bool)
    rim[8:-8, 8:-8] = False
    assert np.abs(eta[rim] - 2.0).max() < 1e-11   # nothing has reached the walls
    assert abs(h[I].sum() - mass0) / mass0 < 1e-12


@pytest.mark.slow
@pytest.mark.parametrize("bump", [True, False])
def test_lake_4096_wetdry_as_stated_stepwise(hip, bump):
    """Config 2 exactly as stated: 4096^2 on [0,10]^2, 32 seeded Gaussian bumps, lake at rest
    eta = 0.5 (h = max(0, eta - B), 23% of the cells dry) with and without the 0.05 bump,
    transmissive walls, CFL 0.4 dt.  The HIP path and the oracle advance one step at a time:
    the full arrays are bitwise equal after every step, until the step at which the
    reference scheme produces a non-finite state (omp_kernels.hpp:309-331: at wet/dry
    fronts MUSCL on eta gives dry cells a face depth and the hydrostatic correction gives
    them momentum, so u = hu/h blows up once h is just above the 1e-8 floor).  At that same
    step the HIP path reports SWE_ERR_NONFINITE and leaves the arrays the reference leaves;
    so does one call of 200 steps from the initial state (first-bad-stage stop)."""
    c = W.lake_wet_dry(4096, seed=2, bump=bump)
    gr, h, hu, hv, B = oracle_state(c)
    dt, _ = oracle.stable_dt(gr, h, hu, hv, 0.4)
    failed_at = None
    with gpu_solver(hip, c) as s:
        for step in range(1, 201):
            rc, _, _ = oracle.run(gr, h, hu, hv, B, 1, dt=dt)
            if rc != 0:
                with pytest.raises(hip.SweError) as e:
                    s.step(dt, 1)
                assert e.value.code == hip.SWE_ERR_NONFINITE
                # stopped where the reference throws, with its arrays (interior)
                got = s.download()
                for comp, g_, w_ in zip(("h", "hu", "hv"), got, (h, hu, hv)):
                    assert_bitwise(g_[2:-2, 2:-2], w_[2:-2, 2:-2],
                                   f"config 2 (bump={bump}) stopped at step {step} {comp}",
                                   nan_any=True)
                failed_at = step
                break
            s.step(dt, 1)
            got = s.download()
            for comp, g_, w_ in zip(("h", "hu", "hv"), got, (h, hu, hv)):
                assert_bitwise(g_, w_, f"config 2 (bump={bump}) step {step} {comp}")
    assert failed_at is not None, "the reference scheme no longer diverges within 200 steps"
    with gpu_solver(hip, c) as s:
        with pytest.raises(hip.SweError) as e:
            s.step(dt, 200)
        assert e.value.code == hip.SWE_ERR_NONFINITE
        status = s.last_status()
        assert status[0] == failed_at
        got = s.download()
    for comp, g_, w_ in zip(("h", "hu", "hv"), got, (h, hu, hv)):
        assert_bitwise(g_[2:-2, 2:-2], w_[2:-2, 2:-2],
                       f"config 2 (bump={bump}) one 200-step call {comp}", nan_any=True)
    print(f"config 2 (bump={bump}): bitwise for {failed_at - 1} steps, non-finite at step "
          f"{failed_at} in both; one 200-step call stops there at stage {status[1]}")


@pytest.mark.slow
def test_dam_break_8192_full_run_properties(hip):
    """Config 3 at its real size for all its 1000 steps (CFL 0.45 dt, reflecting walls):
    the state stays finite, the depth stays between the two initial levels up to the wave
    overshoot (no floor), and the walls conserve mass to round-off."""
    c = W.dam_break(8192)
    I = (slice(2, 8194), slice(2, 8194))
    mass0 = c["h"][I].sum()
    with gpu_solver(hip, c) as s:
        dt, _ = s.stable_dt(0.45)
        s.step(dt, 1000)
        h, hu, hv = s.download()
    del c
    assert np.isfinite(h).all() and np.isfinite(hu).all() and np.isfinite(hv).all()
    assert 0.9 < h[I].min() and h[I].max() < 2.1
    assert abs(h[I].sum() - mass0) / mass0 < 1e-12


@pytest.mark.slow
def test_dam_break_8192_step_subset(hip):
    """Config 3 at its real size (8192^2), 2 steps, bitwise vs the oracle."""
    c = W.dam_break(8192)
    gr, h, hu, hv, B = oracle_state(c)
    dt, _ = oracle.stable_dt(gr, h, hu, hv, 0.45)
    with gpu_solver(hip, c) as s:
        dtg, _ = s.stable_dt(0.45)
        assert dtg == dt
        s.step(dt, 2)
        got = s.download()
    del c
    rc, _, _ = oracle.run(gr, h, hu, hv, B, 2, dt=dt)
    assert rc == 0
    for comp, g_, w_ in zip(("h", "hu", "hv"), got, (h, hu, hv)):
        assert_bitwise(g_, w_, f"dam_break 8192 {comp}")
