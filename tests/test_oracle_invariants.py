"""Pins of the oracle that do not depend on the reference build: closed forms and
invariants of the scheme.  Each is chosen so that a plausible mistake in the restatement
(a dropped term, a wrong sign or index, swapped operands) fails it.

- HLL consistency F(U, U) = F(U) for supercritical states, against the analytic flux
  (hu, hu^2/h + g h^2/2, hu v) -- pins the flux terms and that g enters (also at g = 9.81);
- minmod closed form on hand values;
- still water on a flat bottom: rhs exactly zero; lake at rest over bathymetry (fully wet):
  rhs at round-off (the hydrostatic reconstruction is well balanced);
- reflecting walls: exact mirror symmetry of a symmetric pulse, and mass conserved to
  round-off (the wall mass flux is exactly zero);
- x <-> y transpose: swapping the axes (and hu <-> hv, dx <-> dy) transposes the result
  bitwise, because hll_x and hll_y are mirror images;
- compute_stable_dt at rest: cfl*min(dx,dy)/sqrt(g*h_max);
- transmissive ghost fill (no reference counterpart): ghosts equal the nearest interior cell.
"""
import math

import numpy as np
import pytest

import oracle
from swe_hip import workloads as W


def same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64),
                          np.ascontiguousarray(b).view(np.uint64))


@pytest.mark.parametrize("a,b,want", [(1.0, 2.0, 1.0), (-3.0, -2.0, -2.0), (1.0, -1.0, 0.0),
                                      (0.0, 5.0, 0.0), (2.0, 2.0, 2.0), (-0.5, -4.0, -0.5),
                                      (1e-200, 1e-200, 0.0)])
def test_minmod_closed_form(a, b, want):
    # the last case: the product underflows to 0, so the reference returns 0
    assert oracle.minmod(a, b) == want


@pytest.mark.parametrize("g", [oracle.G_REF, 9.81])
@pytest.mark.parametrize("direction", ["x", "y"])
def test_hll_consistency_supercritical(g, direction):
    rng_u = W.uniform(3, 50, 1, 0.5, 2.0)
    for k, h in enumerate(rng_u):
        c = math.sqrt(g * h)
        un = (2.0 + k * 0.1) * c        # supercritical to the right: sL >= 0
        ut = 0.3 * (k % 5 - 2)
        if direction == "x":
            out = oracle.hll("x", h, h * un, h * ut, h, h * un, h * ut, g=g)
            fn, ft = out[1], out[2]
        else:
            out = oracle.hll("y", h, h * ut, h * un, h, h * ut, h * un, g=g)
            fn, ft = out[2], out[1]
        assert out[0] == pytest.approx(h * un, rel=1e-15)
        assert fn == pytest.approx(h * un * un + 0.5 * g * h * h, rel=1e-14)
        assert ft == pytest.approx(h * un * ut, rel=1e-14, abs=1e-15)


def test_hll_subcritical_symmetric_states_flux_mass_zero():
    # mirror states (h, u) | (h, -u): the HLL mass flux vanishes exactly
    for h, u in [(1.0, 0.3), (0.2, -0.1), (2.0, 1.5)]:
        f = oracle.hll("x", h, h * u, 0.0, h, -h * u, 0.0)
        assert f[0] == 0.0


def test_still_water_flat_rhs_is_exactly_zero():
    c = W.lake_wet_dry(40, bump=False, eta0=1.0, nbumps=0)
    gr = oracle.grid(40, 40, 2, c["dx"], c["dy"], bc=0)
    B = c["B"].copy()
    oracle.apply_bc(gr, B)
    r = oracle.rhs(gr, c["h"].copy(), c["hu"].copy(), c["hv"].copy(), B)
    for a in r:
        assert not a.any()


def test_lake_at_rest_wet_is_balanced():
    c = W.lake_wet_dry(64, seed=2, bump=False, eta0=2.0)
    gr = oracle.grid(64, 64, 2, c["dx"], c["dy"], bc=1)
    B = c["B"].copy()
    oracle.apply_bc(gr, B)
    r = oracle.rhs(gr, c["h"].copy(), c["hu"].copy(), c["hv"].copy(), B)
    assert max(np.abs(a).max() for a in r) < 1e-12
    h, hu, hv = c["h"].copy(), c["hu"].copy(), c["hv"].copy()
    dt, _ = oracle.stable_dt(gr, h, hu, hv, 0.4)
    rc, _, _ = oracle.run(gr, h, hu, hv, B, 50, dt=dt)
    assert rc == 0
    I = (slice(2, 66), slice(2, 66))
    assert np.abs(h[I] + B[I] - 2.0).max() < 1e-12
    assert max(np.abs(hu[I]).max(), np.abs(hv[I]).max()) < 1e-11


def test_mirror_symmetry_and_mass_reflecting():
    c = W.gauss_pulse(48, g=9.81)
    gr = oracle.grid(48, 48, 2, c["dx"], c["dy"], g=9.81, bc=0)
    h, hu, hv, B = c["h"].copy(), c["hu"].copy(), c["hv"].copy(), c["B"].copy()
    m0 = math.fsum(h[2:50, 2:50].ravel())
    dt, _ = oracle.stable_dt(gr, h, hu, hv, 0.4)
    rc, _, _ = oracle.run(gr, h, hu, hv, B, 40, dt=dt)
    assert rc == 0
    I = (slice(2, 50), slice(2, 50))
    hi = h[I]
    # x mirror: h(i) = h(N-1-i), hu(i) = -hu(N-1-i)
    assert np.abs(hi - hi[::-1, :]).max() <= 1e-14
    assert np.abs(hu[I] + hu[I][::-1, :]).max() <= 1e-14
    m1 = math.fsum(hi.ravel())
    assert abs(m1 - m0) / m0 < 1e-13


@pytest.mark.parametrize("steps", [0, 3])
def test_transpose_symmetry(steps):
    c = W.random_state(23, 31, seed=8)
    dx, dy = c["dx"], c["dy"]
    g1 = oracle.grid(23, 31, 2, dx, dy)
    g2 = oracle.grid(31, 23, 2, dy, dx)
    B = c["B"].copy()
    oracle.apply_bc(g1, B)
    h, hu, hv = c["h"].copy(), c["hu"].copy(), c["hv"].copy()
    oracle.apply_bc(g1, h, hu, hv)
    hT, huT, hvT, BT = (np.ascontiguousarray(a.T) for a in (h, hv, hu, B))
    if steps == 0:
        r1 = oracle.rhs(g1, h, hu, hv, B)
        r2 = oracle.rhs(g2, hT, huT, hvT, BT)
        pairs = [(r1[0], r2[0]), (r1[1], r2[2]), (r1[2], r2[1])]
        for a, b in pairs:
            # interior only: ghosts of rhs are 0 on both sides anyway
            assert same(np.ascontiguousarray(a.T), b)
    else:
        dt, _ = oracle.stable_dt(g1, h, hu, hv, 0.2)
        dtT, _ = oracle.stable_dt(g2, hT, huT, hvT, 0.2)
        assert dt == dtT
        oracle.run(g1, h, hu, hv, B, steps, dt=dt)
        oracle.run(g2, hT, huT, hvT, BT, steps, dt=dt)
        I = (slice(2, 25), slice(2, 33))
        for a, b in [(h, hT), (hu, hvT), (hv, huT)]:
            assert same(np.ascontiguousarray(a[I].T), b[I[1], I[0]])


def test_stable_dt_closed_form_at_rest():
    for g in (oracle.G_REF, 9.81):
        c = W.lake_wet_dry(30, seed=4, bump=False, eta0=2.0, g=g)
        gr = oracle.grid(30, 30, 2, c["dx"], c["dy"], g=g)
        dt, ms = oracle.stable_dt(gr, c["h"], c["hu"], c["hv"], 0.45)
        hmax = c["h"][2:32, 2:32].max()
        assert ms == math.sqrt(g * hmax)
        assert dt == 0.45 * min(c["dx"], c["dy"]) / math.sqrt(g * hmax)


def test_stable_dt_dry_returns_sentinel():
    gr = oracle.grid(8, 8, 2, 0.1, 0.1)
    z = np.zeros((12, 12))
    dt, ms = oracle.stable_dt(gr, z, z, z, 0.4)
    assert dt == 1e100 and ms == 0.0


def test_positivity_floor():
    gr = oracle.grid(6, 5, 2, 0.1, 0.1)
    h = np.full((10, 9), 0.5)
    hu = np.full((10, 9), 0.2)
    hv = np.full((10, 9), -0.1)
    h[3, 4] = 5e-9
    h[4, 4] = -1.0
    assert oracle.enforce_positivity(gr, h, hu, hv) == 0
    for i, j in [(3, 4), (4, 4)]:
        assert (h[i, j], hu[i, j], hv[i, j]) == (1e-8, 0.0, 0.0)
    assert h[2, 2] == 0.5 and hu[2, 2] == 0.2
    hv[5, 5] = np.inf
    assert oracle.enforce_positivity(gr, h, hu, hv) == 1


def test_transmissive_ghosts_copy_nearest_interior():
    c = W.random_state(7, 9, seed=3, bc=1)
    gr = oracle.grid(7, 9, 2, c["dx"], c["dy"], bc=1)
    h, hu, hv = c["h"].copy(), c["hu"].copy(), c["hv"].copy()
    oracle.apply_bc(gr, h, hu, hv)
    for a in (h, hu, hv):
        for g in (0, 1):
            assert np.array_equal(a[g, 2:11], a[2, 2:11])
            assert np.array_equal(a[10 - g, 2:11], a[8, 2:11])
            assert np.array_equal(a[:, g], a[:, 2])
            assert np.array_equal(a[:, 12 - g], a[:, 10])


def test_reflecting_ghosts_mirror_with_sign_flips():
    c = W.random_state(7, 9, seed=4)
    gr = oracle.grid(7, 9, 2, c["dx"], c["dy"], bc=0)
    h, hu, hv = c["h"].copy(), c["hu"].copy(), c["hv"].copy()
    oracle.apply_bc(gr, h, hu, hv)
    for g in (0, 1):
        src = 3 - g  # ghost row g mirrors interior row 2ng-1-g
        assert np.array_equal(h[g, 2:11], h[src, 2:11])
        assert np.array_equal(hu[g, 2:11], -hu[src, 2:11])
        assert np.array_equal(hv[g, 2:11], hv[src, 2:11])
        assert np.array_equal(hv[2:9, g], -hv[2:9, 3 - g])
        assert np.array_equal(hu[2:9, g], hu[2:9, 3 - g])


def test_numpy_ghost_fill_matches_oracle():
    for bc in (0, 1):
        c = W.random_state(9, 6, seed=5, bc=bc)
        gr = oracle.grid(9, 6, 2, c["dx"], c["dy"], bc=bc)
        a = c["B"].copy()
        b = c["B"].copy()
        oracle.apply_bc(gr, a)
        W.fill_ghosts(b, 2, bc)
        assert same(a, b)


# ---- where a run stops: the failing stage and the state it leaves ----------------------
# Cases (workload, CFL for a fixed dt from the initial state) whose run goes non-finite in
# each of the three stages; the reference throws there and leaves U^n untouched after a
# stage-1 or stage-2 failure, the floored stage-3 output after a stage-3 failure.
STOP_CASES = {
    "stage1": (lambda: W.random_state(16, 19, seed=0), 3.0, 1, 4),
    "stage2": (lambda: W.random_state(16, 19, seed=1, wet_dry=True), 0.4, 2, 4),
    "stage3": (lambda: W.random_state(16, 19, seed=0), 1.5, 3, 4),
    "lake48_stage3": (lambda: W.lake_wet_dry(48, seed=1), 0.4, 3, 4),
}


def stop_case(name):
    make, cfl, stage, step = STOP_CASES[name]
    c = make()
    gr = oracle.grid(c["nx"], c["ny"], 2, c["dx"], c["dy"], g=c["g"], bc=c["bc"])
    B = c["B"].copy()
    oracle.apply_bc(gr, B)
    U = [c[k].copy() for k in ("h", "hu", "hv")]
    dt, _ = oracle.stable_dt(gr, *U, cfl)
    return c, gr, B, U, dt, stage, step


def rk3_by_stage(gr, U, B, dt):
    """One rk3_step assembled from the oracle's primitives -- rhs (after its ghost fill),
    the stage combinations of omp_solver.hpp:262-266, :288-292, :314-320 and
    enforce_positivity -- in numpy.  Returns the first stage whose output is non-finite
    (U then as the reference leaves it) or 0."""
    r = oracle.rhs(gr, *U, B)
    U1 = [x + dt * y for x, y in zip(U, r)]
    if oracle.enforce_positivity(gr, *U1):
        return 1
    r = oracle.rhs(gr, *U1, B)
    U2 = [0.75 * x + 0.25 * (y + dt * z) for x, y, z in zip(U, U1, r)]
    if oracle.enforce_positivity(gr, *U2):
        return 2
    r = oracle.rhs(gr, *U2, B)
    for x, y, z in zip(U, U2, r):
        x[...] = (1.0 / 3.0) * x + (2.0 / 3.0) * (y + dt * z)
    if oracle.enforce_positivity(gr, *U):
        return 3
    oracle.apply_bc(gr, *U)
    return 0


@pytest.mark.parametrize("name", sorted(STOP_CASES))
def test_run_stops_at_the_failing_stage(name):
    """oracle_run's failing step and stage, and the state it leaves, against the step
    assembled stage by stage: the stage a run reports is the one whose output first holds a
    non-finite value, U^n stays untouched unless that stage is 3, and the arrays match
    bitwise (the stage-3 case over the whole array: its ghosts are those of U^n)."""
    c, gr, B, U, dt, stage, step = stop_case(name)
    V = [x.copy() for x in U]
    rc, done, _, st = oracle.run_status(gr, *U, B, 20, dt=dt)
    assert (rc, done + 1, st) == (1, step, stage)
    for k in range(1, 21):
        before = [x.copy() for x in V]
        s = rk3_by_stage(gr, V, B, dt)
        if s:
            break
    assert (k, s) == (step, stage)
    I = (slice(2, -2), slice(2, -2))
    for x, y, b in zip(U, V, before):
        assert same(x, y)
        assert same(x[I], b[I]) == (stage != 3)
