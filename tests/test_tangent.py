"""The CPU dual-number tangent (oracle/swe_tangent.cpp), checker of the GPU adjoint, and the
host-side binomial checkpoint schedule.

The tangent is pinned by (a) its primal being bitwise the oracle's fixed-dt run, and (b)
central finite differences of that run on rough random states (no exact ties, so the map
is smooth at the sample point): agreement to ~1e-9 relative at several step sizes.  At
exact ties (a flat lake at rest: neighbouring differences exactly 0) the discrete map is
not differentiable -- minmod's directional derivative at (0, 0) is minmod(da, db), not
linear in the direction -- which test_flat_lake_is_not_differentiable documents.
"""
import math

import numpy as np
import pytest

import oracle
from swe_hip import workloads as W


def same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64),
                          np.ascontiguousarray(b).view(np.uint64))


def interior(shape, rng, ng=2):
    a = np.zeros(shape)
    a[ng:-ng, ng:-ng] = rng.standard_normal((shape[0] - 2 * ng, shape[1] - 2 * ng))
    return a


def prepared(case, bc, cfl=0.3):
    gr = oracle.grid(case["nx"], case["ny"], 2, case["dx"], case["dy"], g=case["g"], bc=bc)
    B = case["B"].copy()
    oracle.apply_bc(gr, B)
    dt, _ = oracle.stable_dt(gr, case["h"], case["hu"], case["hv"], cfl)
    return gr, B, dt


@pytest.mark.parametrize("bc", [0, 1])
def test_tangent_primal_is_the_oracle(bc):
    case = W.random_state(21, 17, seed=3, bathy=True, bc=bc)
    gr, B, dt = prepared(case, bc)
    a = [case[k].copy() for k in ("h", "hu", "hv")]
    assert oracle.tangent_primal(gr, *a, B, np.full(6, dt)) == 0
    b = [case[k].copy() for k in ("h", "hu", "hv")]
    rc, _, _ = oracle.run(gr, *b, B, 6, dt=dt)
    assert rc == 0
    for x, y in zip(a, b):
        assert same(x, y)


@pytest.mark.parametrize("bc,nsteps", [(0, 1), (0, 5), (1, 5)])
def test_tangent_matches_central_differences(bc, nsteps):
    case = W.random_state(16, 13, seed=5, bathy=True, bc=bc)
    gr, B, dt = prepared(case, bc, 0.4)
    dts = np.full(nsteps, dt)
    rng = np.random.default_rng(nsteps + 10 * bc)
    shape = case["h"].shape
    v = [1e-2 * interior(shape, rng) for _ in range(4)]
    x = [case[k].copy() for k in ("h", "hu", "hv")]
    d = [a.copy() for a in v[:3]]
    assert oracle.tangent_run(gr, *x, B, *d, v[3], dts) == 0
    for eps in (1e-4, 1e-5):
        outs = []
        for sgn in (1.0, -1.0):
            y = [np.ascontiguousarray(case[q] + sgn * eps * vv)
                 for q, vv in zip(("h", "hu", "hv"), v[:3])]
            Bp = B + sgn * eps * v[3]
            oracle.apply_bc(gr, Bp)
            assert oracle.tangent_primal(gr, *y, Bp, dts) == 0
            outs.append(y)
        fd = [(p - m) / (2 * eps) for p, m in zip(*outs)]
        err = max(np.abs(f - t).max() for f, t in zip(fd, d)) / max(np.abs(t).max() for t in d)
        assert err < 1e-8, (eps, err)


def test_tangent_is_linear_in_the_direction():
    case = W.random_state(14, 15, seed=8, bathy=True)
    gr, B, dt = prepared(case, 0)
    dts = np.full(3, dt)
    rng = np.random.default_rng(2)
    shape = case["h"].shape
    v1 = [interior(shape, rng) for _ in range(4)]
    v2 = [interior(shape, rng) for _ in range(4)]

    def tan(v):
        x = [case[k].copy() for k in ("h", "hu", "hv")]
        d = [a.copy() for a in v[:3]]
        assert oracle.tangent_run(gr, *x, B, *d, v[3], dts) == 0
        return d

    t1, t2 = tan(v1), tan(v2)
    t12 = tan([2.0 * a + 3.0 * b for a, b in zip(v1, v2)])
    for a, b, c in zip(t1, t2, t12):
        assert np.allclose(c, 2.0 * a + 3.0 * b, rtol=1e-12, atol=1e-12 * np.abs(c).max())


def test_flat_lake_is_not_differentiable():
    """Exact ties make central differences disagree with any linear derivative: a lake at
    rest (eta = 1.0 exactly away from the bump) perturbed along a smooth direction."""
    case = W.adjoint_case(24, steps=10, background=False)
    gr, B, dt = prepared(case, 0, 0.4)
    dts = np.full(10, dt)
    shape = case["h"].shape
    x_ = (np.arange(24) + 0.5) / 24
    deta = np.zeros(shape)
    deta[2:-2, 2:-2] = np.outer(np.sin(2 * np.pi * x_), np.cos(2 * np.pi * x_))
    zero = np.zeros(shape)
    x = [case[k].copy() for k in ("h", "hu", "hv")]
    d = [deta.copy(), zero.copy(), zero.copy()]
    assert oracle.tangent_run(gr, *x, B, *d, zero, dts) == 0
    eps = 1e-7
    outs = []
    for sgn in (1.0, -1.0):
        y = [np.ascontiguousarray(case["h"] + sgn * eps * deta), case["hu"].copy(),
             case["hv"].copy()]
        assert oracle.tangent_primal(gr, *y, B, dts) == 0
        outs.append(y)
    fd = (outs[0][0] - outs[1][0]) / (2 * eps)
    assert np.abs(fd - d[0]).max() > 1e-3 * np.abs(d[0]).max()


# ---- checkpoint schedule (host only) -----------------------------------------------------

def binomial_optimum(n, s):
    """Griewank's minimal number of forward steps to reverse n steps with s stored states."""
    r = 0
    while math.comb(s + r, s) < n:
        r += 1
    return r * n - math.comb(s + r, r - 1) if r > 0 else 0


def brute_force(n, f, memo={}):
    """Minimal advances by exhaustive search over the first checkpoint position."""
    if (n, f) in memo:
        return memo[(n, f)]
    if n == 1:
        v = 0
    elif f == 0:
        v = n * (n - 1) // 2
    else:
        v = min(k + brute_force(n - k, f - 1) + brute_force(k, f) for k in range(1, n))
    memo[(n, f)] = v
    return v


def test_schedule_is_optimal():
    from swe_hip import adjoint_schedule

    for n in range(1, 60):
        for s in range(1, 7):
            adv, reps, acts = adjoint_schedule(n, s)
            assert adv == brute_force(n, s - 1) == binomial_optimum(n, s), (n, s)
    for n, s in [(2000, 64), (2000, 16), (2000, 2000), (12345, 37)]:
        adv, reps, _ = adjoint_schedule(n, s)
        assert adv == binomial_optimum(n, s)
    # store-all: every step advanced exactly once (the first sweep)
    assert adjoint_schedule(2000, 2000)[:2] == (1999, 1)


def test_record_plan_is_the_best_that_fits():
    """The budget plan (the last K steps' stage records, a checkpoint every m steps before
    them, two segment buffers of m records when there are two segments or more) maximises
    K within the budget: brute force over every (K, m) on small cases, in units of one
    array (n_total = 1 cell, 8 bytes)."""
    from swe_hip import adjoint_plan_budget

    def need(n, K, m):
        nck = -(-(n - K) // m)
        return 9 * max(K, 2 * m if nck >= 2 else m) + 3 * nck

    for n in range(1, 40):
        for arrays in list(range(3, 12 * n, 7)) + [9 * n, 9 * n + 5]:
            K, m, states = adjoint_plan_budget(n, 1, 8.0 * arrays)
            if 9 * n <= arrays:
                assert (K, states) == (n, n), (n, arrays)
                continue
            best = max((k for k in range(1, n) for mm in range(1, n + 1)
                        if need(n, k, mm) <= arrays), default=0)
            assert K == best, (n, arrays, K, best)
            if K:
                assert need(n, K, m) <= arrays and states == K + -(-(n - K) // m)
            else:  # binomial fallback: as many 3-array states as fit
                assert states == min(n, arrays // 3)
    # config 5 at a quarter of the store-every-step memory: about 500 of 2000 steps recorded
    n_total = 1028 * 1028
    K, m, _ = adjoint_plan_budget(2000, n_total, 0.25 * 2000 * 9 * n_total * 8)
    assert 480 <= K <= 500 and m >= 1
