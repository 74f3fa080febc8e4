"""swe_hip_step_adaptive_noalloc (the entry point that allocates no device memory) against
swe_hip_step_adaptive: the same states and dt histories bit for bit, on two calls in a row
on one GPU (the second on a smaller grid, so the shared scratch holds the first call's
maxima), and the size limit of its fixed scratch reported as an error."""
import numpy as np
import pytest

from swe_hip import workloads as W

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def run_pair(hip, torch, case, nsteps, cfl=0.4):
    with hip.HipSolver.from_case(case) as s:
        s.upload(case["h"], case["hu"], case["hv"], case["B"])
        want_dt = s.step_adaptive(cfl, nsteps)
        want = s.download()
    hist = torch.full((nsteps,), -1.0, dtype=torch.float64, device="cuda:0")
    with hip.HipSolver.from_case(case) as s:
        s.upload(case["h"], case["hu"], case["hv"], case["B"])
        s.step_adaptive_noalloc(cfl, nsteps, hist)
        got = s.download()
    torch.cuda.synchronize()
    return got, want, hist.cpu().numpy(), want_dt


def test_noalloc_matches_step_adaptive_twice(hip):
    torch = pytest.importorskip("torch")
    big = W.lake_wet_dry(150, seed=2, bump=True, eta0=2.0)       # 3 strips, transmissive
    small = W.random_state(37, 61, seed=5, bathy=True, bc=0)      # reflecting, 2 strips
    for case, nsteps in ((big, 25), (small, 7)):
        got, want, hist, want_dt = run_pair(hip, torch, case, nsteps)
        for g_, w_, name in zip(got, want, ("h", "hu", "hv")):
            assert np.array_equal(bits(g_), bits(w_)), (case["nx"], name)
        assert np.array_equal(bits(hist), bits(want_dt)), case["nx"]


def test_noalloc_without_history_and_size_limit(hip):
    torch = pytest.importorskip("torch")  # noqa: F841 (device context)
    case = W.random_state(40, 64, seed=9, bathy=True, bc=1)
    with hip.HipSolver.from_case(case) as s:
        s.upload(case["h"], case["hu"], case["hv"], case["B"])
        want_dt = s.step_adaptive(0.3, 5)
        want = s.download()
    with hip.HipSolver.from_case(case) as s:
        s.upload(case["h"], case["hu"], case["hv"], case["B"])
        s.step_adaptive_noalloc(0.3, 5)
        got = s.download()
    assert len(want_dt) == 5
    for g_, w_ in zip(got, want):
        assert np.array_equal(bits(g_), bits(w_))
    # 1920 columns (32 strips = 8 workgroups wide) x 8200 rows at 1 row per wave: 65600
    # workgroups, beyond the 65536 per-block maxima of the scratch
    with hip.HipSolver(8200, 1920, 1.0 / 8200, 1.0 / 1920) as s:
        s.tile_rows = 1
        with pytest.raises(hip.SweError) as e:
            s.step_adaptive_noalloc(0.4, 1)
        assert e.value.code == hip.SWE_ERR_INVALID and "65536" in str(e.value)
