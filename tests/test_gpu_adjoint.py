"""Discrete adjoint on the GPU (swe_hip_adj_* through the C ABI).

The adjoint is pinned by the CPU dual-number tangent of the serial solver
(oracle/swe_tangent.cpp), which uses the same one-sided derivative conventions: for random
output weights lam and input directions v, <lam, J v> (tangent) must equal <J^T lam, v>
(adjoint) to rounding.  The forward pass under the adjoint object must equal the plain
solver bitwise, and the gradient must not depend on the number of stored checkpoints.
"""
import numpy as np
import pytest

import oracle
from swe_hip import workloads as W

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def same(a, b):
    return np.array_equal(bits(a), bits(b))


def interior(shape, rng, ng=2, scale=1.0):
    a = np.zeros(shape)
    a[ng:-ng, ng:-ng] = scale * rng.standard_normal((shape[0] - 2 * ng, shape[1] - 2 * ng))
    return a


def dry_dam(nx, ny):
    """Dam break onto a dry bed (h = 1e-8, the floor state) over a flat bottom."""
    c = W._case(nx, ny, 1.0, 1.0, 9.80665, 0, 0.3, 1, 2, name=f"drydam_{nx}x{ny}")
    x = (np.arange(nx) + 0.5) / nx
    c["h"][2:-2, 2:-2] = np.where(x[:, None] < 0.4, 1.0, 1e-8) * np.ones((1, ny))
    W.fill_ghosts(c["h"], 2, 0)
    return c


def setup(case, bc):
    gr = oracle.grid(case["nx"], case["ny"], 2, case["dx"], case["dy"], g=case["g"], bc=bc)
    B = case["B"].copy()
    oracle.apply_bc(gr, B)
    dt, _ = oracle.stable_dt(gr, case["h"], case["hu"], case["hv"], 0.3)
    return gr, B, dt


def gpu_gradient(hip, case, bc, nsteps, dt, lam, nsnap=2, launch=None, budget=None):
    """launch: (block_waves, rows, y_rows) for HipAdjoint.set_launch, None = default;
    budget: bytes for the budget storage plan (instead of nsnap)."""
    c = dict(case)
    c["bc"] = bc
    with hip.HipSolver.from_case(c) as s:
        s.upload(case["h"], case["hu"], case["hv"], case["B"])
        with hip.HipAdjoint(s, nsteps, nsnap, budget_bytes=budget) as adj:
            if launch:
                adj.set_launch(*launch)
            adj.forward(dt=dt)
            final = s.download()
            grads = adj.backward(*lam)
    return final, grads


def mild_state(nx, ny, seed, bc=0, bathy_scale=0.15):
    """Rough random state whose bottom variation stays well below the depth, so every
    hydrostatic face depth is bounded away from 0 and the Jacobian is well conditioned."""
    case = W.random_state(nx, ny, seed=seed, bathy=True, bc=bc)
    case["B"] = np.ascontiguousarray(case["B"] * bathy_scale)
    return case


CASES = [
    # (nx, ny, nsteps, bc, launch, bathy_scale, rel_tol): multi-strip (ny > 58), blocks
    # shorter than the grid, single-row waves, blocks with idle waves, tiny grids, and the
    # two-kernel path (block_waves 0).  The last case keeps the full random bottom: nearly
    # dry hydrostatic faces make the Jacobian huge (|J| ~ 1e40), so only its rounding-level
    # agreement bound is looser.
    (20, 17, 1, 0, None, 0.15, 1e-10),
    (40, 70, 3, 0, (8, 1, 0), 0.15, 1e-10),
    (33, 130, 2, 1, (4, 3, 0), 0.15, 1e-10),
    (70, 61, 4, 0, (0, 16, 8), 0.15, 1e-10),
    (9, 7, 3, 1, (8, 4, 0), 0.15, 1e-10),
    (12, 120, 2, 0, (4, 1, 0), 0.15, 1e-10),
    (70, 61, 4, 0, (8, 2, 0), 1.0, 1e-8),
]


@pytest.mark.parametrize("nx,ny,nsteps,bc,launch,bscale,tol", CASES)
def test_adjoint_matches_tangent(hip, nx, ny, nsteps, bc, launch, bscale, tol):
    case = mild_state(nx, ny, nx * 7 + ny, bc, bscale)
    gr, B, dt = setup(case, bc)
    rng = np.random.default_rng(nx + 100 * ny)
    shape = case["h"].shape
    lam = [interior(shape, rng) for _ in range(3)]
    final, (gh, ghu, ghv, gb) = gpu_gradient(hip, case, bc, nsteps, dt, lam, launch=launch)

    # forward under the adjoint object == the oracle's run, bitwise
    ref = [case[k].copy() for k in ("h", "hu", "hv")]
    rc, _, _ = oracle.run(gr, *ref, B, nsteps, dt=dt)
    assert rc == 0
    for g_, w_ in zip(final, ref):
        assert same(g_, w_)

    # ghost entries of the gradients are zero
    for g_ in (gh, ghu, ghv, gb):
        assert not g_[:2].any() and not g_[-2:].any() and not g_[:, :2].any() and not g_[:, -2:].any()

    dts = np.full(nsteps, dt)
    grads = (gh, ghu, ghv, gb)
    errs = []
    # one direction per input (h, hu, hv, B), then a mixed one
    for k in range(5):
        v = [interior(shape, rng) if (k == 4 or i == k) else np.zeros(shape) for i in range(4)]
        x = [case[q].copy() for q in ("h", "hu", "hv")]
        d = [a.copy() for a in v[:3]]
        assert oracle.tangent_run(gr, *x, B, *d, v[3], dts) == 0
        lhs_terms = [lam[i] * d[i] for i in range(3)]
        lhs = sum(t.sum() for t in lhs_terms)
        rhs = sum((g_ * v_).sum() for g_, v_ in zip(grads, v))
        scale = sum(np.abs(t).sum() for t in lhs_terms)
        errs.append(abs(lhs - rhs) / scale)
    print("relative dot-product errors (h, hu, hv, B, mixed):", errs)
    assert max(errs) <= tol, errs


def test_adjoint_floor_mask_matches_tangent(hip):
    case = dry_dam(40, 66)
    gr, B, dt = setup(case, 0)
    rng = np.random.default_rng(5)
    shape = case["h"].shape
    lam = [interior(shape, rng) for _ in range(3)]
    nsteps = 4
    _, (gh, ghu, ghv, gb) = gpu_gradient(hip, case, 0, nsteps, dt, lam)
    dts = np.full(nsteps, dt)
    for k in range(2):
        v = [interior(shape, rng) for _ in range(4)]
        x = [case[q].copy() for q in ("h", "hu", "hv")]
        d = [a.copy() for a in v[:3]]
        assert oracle.tangent_run(gr, *x, B, *d, v[3], dts) == 0
        floored = (x[0] == 1e-8) & (x[1] == 0) & (x[2] == 0)
        assert floored[2:-2, 2:-2].sum() > 0  # the case exercises the floor mask
        lhs = sum((lam[i] * d[i]).sum() for i in range(3))
        rhs = (gh * v[0]).sum() + (ghu * v[1]).sum() + (ghv * v[2]).sum() + (gb * v[3]).sum()
        scale = sum(np.abs(lam[i] * d[i]).sum() for i in range(3))
        assert abs(lhs - rhs) <= 1e-11 * scale, (lhs, rhs)


def test_gradient_independent_of_checkpoints(hip):
    case = mild_state(30, 64, 11)
    _, _, dt = setup(case, 0)
    rng = np.random.default_rng(3)
    lam = [interior(case["h"].shape, rng) for _ in range(3)]
    ref = None
    for nsnap in (9, 1, 2, 3, 5, -1):
        _, g = gpu_gradient(hip, case, 0, 9, dt, lam, nsnap=nsnap)
        if ref is None:
            ref = g
        else:
            for a, b in zip(g, ref):
                assert same(a, b), nsnap
    # budget plans: records of the last K steps plus checkpoints every m steps, down to the
    # binomial fallback (budgets in arrays of one field)
    arr = case["h"].size * 8
    plans = set()
    for arrays in (81, 75, 66, 57, 48, 45, 39, 27, 20, 9, 3):
        with hip.HipSolver.from_case(case) as s, \
                hip.HipAdjoint(s, 9, budget_bytes=arrays * arr) as adj:
            p = adj.plan()
        plans.add((p["record_steps"], p["ckpt_every"]))
        assert p["store_bytes"] <= arrays * arr
        _, g = gpu_gradient(hip, case, 0, 9, dt, lam, budget=arrays * arr)
        for a, b in zip(g, ref):
            assert same(a, b), (arrays, p)
    # store-all, records with one segment, records with 2+ segments (overlapped recompute
    # on the second stream) and binomial checkpoints
    assert len(plans) >= 7 and (9, 1) in plans and any(k == 0 for k, _ in plans)
    assert any(k and -(-(9 - k) // m) >= 2 for k, m in plans)


def test_binomial_overlap_long_schedule(hip):
    """The binomial backward issues its advances and stage states on a second stream, one
    reversal ahead, with double-buffered stage states and work state and waits on the last
    reader of every slot: over 40 steps with 2, 4 and 7 slots the schedule reuses slots and
    alternates the work buffers many times; the gradient must be the bits of the
    store-every-step plan (a different executor on one stream)."""
    case = mild_state(30, 64, 13)
    _, _, dt = setup(case, 0)
    rng = np.random.default_rng(5)
    lam = [interior(case["h"].shape, rng) for _ in range(3)]
    _, ref = gpu_gradient(hip, case, 0, 40, dt, lam, nsnap=-1)
    for nsnap in (2, 4, 7):
        _, g = gpu_gradient(hip, case, 0, 40, dt, lam, nsnap=nsnap)
        for x, y in zip(g, ref):
            assert same(x, y), nsnap


@pytest.mark.parametrize("bc", [0, 1])
def test_gradient_independent_of_launch_shape(hip, bc):
    """The fused kernel (x faces through LDS, blocks overlapping by 3 faces) and the
    two-kernel path (x faces through device memory) sum the same terms in the same order:
    every launch shape gives the same bits."""
    case = mild_state(50, 61, 12, bc=bc)
    _, _, dt = setup(case, bc)
    rng = np.random.default_rng(4)
    lam = [interior(case["h"].shape, rng) for _ in range(3)]
    _, ref = gpu_gradient(hip, case, bc, 3, dt, lam, launch=(0, 16, 8))
    for launch in ((0, 1, 1), (0, 7, 3), (0, 64, 64), (8, 1, 0), (8, 2, 0), (8, 4, 0),
                   (4, 1, 0), (4, 3, 0), (4, 9, 0), (4, 7, 0), (2, 1, 0), (2, 5, 0)):
        _, g = gpu_gradient(hip, case, bc, 3, dt, lam, launch=launch)
        for a, b in zip(g, ref):
            assert same(a, b), launch


def test_launch_shape_limits(hip):
    case = mild_state(20, 17, 1)
    with hip.HipSolver.from_case(case) as s, hip.HipAdjoint(s, 2, 2) as adj:
        for bad in ((16, 2, 0), (3, 1, 0), (8, 5, 0), (4, 10, 0), (2, 19, 0)):  # W, LDS
            with pytest.raises(hip.SweError):
                adj.set_launch(*bad)
        adj.set_launch(8, 4, 0)
        adj.set_launch(4, 9, 0)
        adj.set_launch(2, 18, 0)


def test_adjoint_adaptive_dt_is_recorded_and_frozen(hip):
    case = W.adjoint_case(48, steps=12)
    with hip.HipSolver.from_case(case) as s:
        s.upload(case["h"], case["hu"], case["hv"], case["B"])
        want_hist = s.step_adaptive(0.4, 12)
        want = s.download()
    with hip.HipSolver.from_case(case) as s:
        s.upload(case["h"], case["hu"], case["hv"], case["B"])
        with hip.HipAdjoint(s, 12, 4) as adj:
            hist = adj.forward(cfl=0.4)
            got = s.download()
    assert same(hist, want_hist)
    for a, b in zip(got, want):
        assert same(a, b)


def test_backward_requires_fresh_forward(hip):
    case = W.random_state(12, 10, seed=1)
    _, _, dt = setup(case, 0)
    lam = [np.zeros(case["h"].shape)] * 3
    with hip.HipSolver.from_case(case) as s:
        s.upload(case["h"], case["hu"], case["hv"], case["B"])
        with hip.HipAdjoint(s, 2, 2) as adj:
            with pytest.raises(hip.SweError):
                adj.backward(*lam)
            adj.forward(dt=dt)
            adj.backward(*lam)
            with pytest.raises(hip.SweError):
                adj.backward(*lam)


def test_torch_autograd_wrapper(hip):
    torch = pytest.importorskip("torch")
    from swe_hip.autograd import swe_rollout

    case = mild_state(24, 70, 21)
    _, _, dt = setup(case, 0)
    rng = np.random.default_rng(8)
    lam = [interior(case["h"].shape, rng) for _ in range(3)]
    final, (gh, ghu, ghv, gb) = gpu_gradient(hip, case, 0, 5, dt, lam, nsnap=3)

    dev = torch.device("cuda:0")
    t = {k: torch.tensor(case[k], dtype=torch.float64, device=dev, requires_grad=True)
         for k in ("h", "hu", "hv", "B")}
    grid = {"nx": case["nx"], "ny": case["ny"], "dx": case["dx"], "dy": case["dy"],
            "g": case["g"], "bc": 0}
    hN, huN, hvN = swe_rollout(t["h"], t["hu"], t["hv"], t["B"], grid, 5, dt=dt, nsnap=3)
    for a, b in zip((hN, huN, hvN), final):
        assert same(a.detach().cpu().numpy(), b)
    L = sum((torch.tensor(l, device=dev) * o).sum() for l, o in zip(lam, (hN, huN, hvN)))
    L.backward()
    for k, want in zip(("h", "hu", "hv", "B"), (gh, ghu, ghv, gb)):
        assert same(t[k].grad.cpu().numpy(), want), k


def test_torch_rollout_without_grad_skips_the_adjoint(hip):
    """Under no_grad, or with no input requiring grad, swe_rollout runs the plain solver
    (no checkpoint store) and returns the same bits as the differentiable path."""
    torch = pytest.importorskip("torch")
    from swe_hip.autograd import swe_rollout

    case = mild_state(24, 70, 21)
    _, _, dt = setup(case, 0)
    dev = torch.device("cuda:0")
    grid = {"nx": case["nx"], "ny": case["ny"], "dx": case["dx"], "dy": case["dy"],
            "g": case["g"], "bc": 0}
    t = {k: torch.tensor(case[k], dtype=torch.float64, device=dev, requires_grad=True)
         for k in ("h", "hu", "hv", "B")}
    want = swe_rollout(t["h"], t["hu"], t["hv"], t["B"], grid, 5, dt=dt, nsnap=3)
    assert all(o.requires_grad for o in want)
    with torch.no_grad():
        got = swe_rollout(t["h"], t["hu"], t["hv"], t["B"], grid, 5, dt=dt, nsnap=3)
    plain = {k: v.detach() for k, v in t.items()}
    got2 = swe_rollout(plain["h"], plain["hu"], plain["hv"], plain["B"], grid, 5, dt=dt)
    for outs in (got, got2):
        assert not any(o.requires_grad for o in outs)
        for a, b in zip(outs, want):
            assert same(a.cpu().numpy(), b.detach().cpu().numpy())
    dts = [dt * 0.5] * 5
    a = swe_rollout(plain["h"], plain["hu"], plain["hv"], plain["B"], grid, 5, dts=dts)
    b = swe_rollout(t["h"], t["hu"], t["hv"], t["B"], grid, 5, dts=dts, nsnap=3)
    for x, y in zip(a, b):
        assert same(x.cpu().numpy(), y.detach().cpu().numpy())


def config5_fd_errors(hip, steps, eps):
    """Relative |FD - adjoint| of the config-5 gauge-misfit loss (1024^2, the first `steps`
    steps) on 8 smooth random directions in (eta0, b): central differences of the GPU
    forward with half-width eps against the adjoint's directional derivative."""
    n = 1024
    case = W.adjoint_case(n, steps=steps)
    rows, cols = W.adjoint_gauges(n)
    with hip.HipSolver.from_case(case) as s:
        s.upload(case["h"], case["hu"], case["hv"], case["B"])
        dt, _ = s.stable_dt(0.4)

    def run(h0, B):
        with hip.HipSolver.from_case(case) as s:
            s.upload(h0, case["hu"], case["hv"], B)
            s.step(dt, steps)
            return s.download()[0][rows, cols]

    target = run(case["h"], case["B"]) * 0.99
    with hip.HipSolver.from_case(case) as s:
        s.upload(case["h"], case["hu"], case["hv"], case["B"])
        with hip.HipAdjoint(s, steps, 0) as adj:
            adj.forward(dt=dt)
            hN = s.download()[0]
            lam = np.zeros_like(hN)
            np.add.at(lam, (rows, cols), 2.0 * (hN[rows, cols] - target))
            z = np.zeros_like(hN)
            gh, _, _, gb = adj.backward(lam, z, z)
    g_eta, g_b = gh, gb - gh  # h0 = eta0 - b (wet everywhere)
    assert (case["h"][2:-2, 2:-2] > 0.1).all()

    L = 10.0
    x = (np.arange(n) + 0.5) * L / n
    errs = []
    for k in range(8):
        r = np.random.default_rng(100 + k)
        dirs = []
        for _ in range(2):
            f = np.zeros((n, n))
            for _ in range(4):
                kx, ky = r.integers(1, 4, 2)
                px, py = r.uniform(0, 2 * np.pi, 2)
                f += r.standard_normal() * np.outer(np.sin(2 * np.pi * kx * x / L + px),
                                                    np.sin(2 * np.pi * ky * x / L + py))
            a = np.zeros_like(case["h"])
            a[2:-2, 2:-2] = f / np.abs(f).max()
            dirs.append(a)
        de, db = dirs[0], 0.3 * dirs[1]
        gv = float((g_eta * de).sum() + (g_b * db).sum())
        lp = float(np.sum((run(case["h"] + eps * (de - db), case["B"] + eps * db) - target) ** 2))
        lm = float(np.sum((run(case["h"] - eps * (de - db), case["B"] - eps * db) - target) ** 2))
        fd = (lp - lm) / (2 * eps)
        errs.append(abs(fd - gv) / abs(gv))
    return errs


@pytest.mark.parametrize("steps,median_bar,max_bar", [(10, 1e-6, 1e-6), (50, 1e-6, 5e-6)])
def test_config5_finite_differences_short_rollouts(hip, steps, median_bar, max_bar):
    """Where central differences resolve the derivative to the 1e-6 bar: the config-5
    workload over its first 10 steps (all 8 directions within 1e-6; measured max 8e-8) and
    50 steps (median within 1e-6; measured 1.3e-7, max 1.5e-6), eps = 1e-8.  Over more
    steps the kink crossings of the piecewise-smooth scheme dominate the FD error (median
    2.7e-6 at 200 steps, 1.5e-5 at 2000)."""
    errs = config5_fd_errors(hip, steps, 1e-8)
    print(f"{steps} steps: relative |FD - adjoint| on 8 directions:", errs)
    assert np.median(errs) <= median_bar and max(errs) <= max_bar, errs


@pytest.mark.slow
def test_config5_finite_differences(hip):
    """1024^2, 2000 steps, gauge-misfit loss: the adjoint's directional derivatives against
    central differences of the GPU forward on 8 smooth random directions in (eta0, b).  The
    discrete map is only piecewise smooth (minmod, wave-speed and branch kinks everywhere),
    so agreement is limited by kink crossings; the bound here is what that allows."""
    errs = config5_fd_errors(hip, 2000, 1e-8)
    print("relative |FD - adjoint| on 8 directions:", errs)
    assert np.median(errs) < 1e-4 and max(errs) < 1e-3, errs


@pytest.mark.slow
def test_config5_matches_tangent_at_128(hip):
    """The config-5 workload at 128^2 over 2000 steps: <lam, J v> from the CPU dual-number
    tangent equals <J^T lam, v> from the GPU adjoint to rounding."""
    n, steps = 128, 2000
    case = W.adjoint_case(n, steps=steps)
    rows, cols = W.adjoint_gauges(n)
    gr, B, _ = setup(case, 0)
    with hip.HipSolver.from_case(case) as s:
        s.upload(case["h"], case["hu"], case["hv"], case["B"])
        dt, _ = s.stable_dt(0.4)
        with hip.HipAdjoint(s, steps, 16) as adj:   # a binomial schedule
            adj.forward(dt=dt)
            hN = s.download()[0]
            lam = np.zeros_like(hN)
            lam[rows, cols] = 1.0
            z = np.zeros_like(hN)
            gh, ghu, ghv, gb = adj.backward(lam, z, z)
    rng = np.random.default_rng(9)
    shape = case["h"].shape
    for k in range(2):
        v = [interior(shape, rng) for _ in range(4)]
        x_ = [case[q].copy() for q in ("h", "hu", "hv")]
        d = [a.copy() for a in v[:3]]
        assert oracle.tangent_run(gr, *x_, B, *d, v[3], np.full(steps, dt)) == 0
        lhs = float(d[0][rows, cols].sum())
        rhs = float((gh * v[0]).sum() + (ghu * v[1]).sum() + (ghv * v[2]).sum() + (gb * v[3]).sum())
        assert abs(lhs - rhs) <= 1e-9 * max(abs(lhs), 1e-300), (lhs, rhs)
