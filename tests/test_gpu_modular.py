"""GPU parity of the modular serial scheme family (swe_modular.hip) against the modular
oracle (oracle/swe_modular.c, itself pinned bitwise to the reference's own modular headers)
and directly against the reference fixtures in tests/golden/modular.npz.

Bar: bitwise equality of every float64 bit (ghosts included after a step)."""
import os

import numpy as np
import pytest

import oracle
from oracle import RECON_MUSCL, RECON_PC, RIEMANN_HLL, RIEMANN_ROE, RIEMANN_RUSANOV

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "modular.npz")
RECONS = {RECON_MUSCL: "muscl", RECON_PC: "pc"}
RIEMANNS = {RIEMANN_HLL: "hll", RIEMANN_RUSANOV: "rusanov", RIEMANN_ROE: "roe"}
# the combinations the reference's SolverAssembly runs (solver_assembly.hpp:315-500)
SUPPORTED = ([(0, r, q) for r in RECONS for q in RIEMANNS] +
             [(1, r, RIEMANN_HLL) for r in RECONS])


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def scheme_of(hip, bathy):
    return hip.SCHEME_SERIAL if bathy else hip.SCHEME_SERIAL_PLAIN


def wet_state(nx, ny, seed, bathy, lx, ly):
    rng = np.random.default_rng(seed)
    gr = oracle.grid(nx, ny, lx=lx, ly=ly)
    shape = (nx + 4, ny + 4)
    B = 0.3 * rng.random(shape) if bathy else np.zeros(shape)
    h = 0.7 + 0.5 * rng.random(shape)
    hu, hv = 0.3 * rng.standard_normal(shape), 0.3 * rng.standard_normal(shape)
    oracle.mod_apply_bc(gr, B)
    oracle.mod_apply_bc(gr, h, hu, hv)
    return gr, h, hu, hv, B


def solver(hip, gr, bathy, recon, riemann, tile_rows=None):
    s = hip.HipSolver(gr.nx, gr.ny, gr.dx, gr.dy, gr.g, hip.BC_REFLECTING)
    s.set_scheme(scheme_of(hip, bathy), recon, riemann)
    if tile_rows:
        s.tile_rows = tile_rows
    return s


@pytest.mark.parametrize("bathy,recon,riemann", SUPPORTED)
def test_rhs_and_steps_match_oracle(hip, bathy, recon, riemann):
    for (nx, ny, tile) in ((17, 13, None), (40, 130, 7), (70, 61, 1)):
        lx, ly = 0.0625 * nx, 0.125 * ny
        gr, h, hu, hv, B = wet_state(nx, ny, nx + ny + 3 * recon + riemann, bathy, lx, ly)
        with solver(hip, gr, bathy, recon, riemann, tile) as s:
            s.upload(h, hu, hv, B)
            got = s.rhs()
            rc, want = oracle.mod_rhs(gr, bathy, recon, riemann, h.copy(), hu.copy(), hv.copy(), B)
            assert rc == 0
            for g_, w_, nm in zip(got, want, ("h", "hu", "hv")):
                assert same(g_, w_), (nx, ny, tile, "rhs", nm)
            dt = 0.1 * min(gr.dx, gr.dy)
            s.upload(h, hu, hv, B)
            s.step(dt, 6)
            got = s.download()
            ref = [a.copy() for a in (h, hu, hv)]
            assert oracle.mod_run(gr, bathy, recon, riemann, *ref, B, 6, dt) == 0
            for g_, w_, nm in zip(got, ref, ("h", "hu", "hv")):
                assert same(g_, w_), (nx, ny, tile, "steps", nm)


@pytest.mark.parametrize("case", ["dry_cells", "tiny_momentum", "huge_momentum"])
@pytest.mark.parametrize("bathy,recon,riemann", SUPPORTED)
def test_range_guard_fallback(hip, bathy, recon, riemann, case):
    """Waves that load a cell outside the modular range guard (a dry cell, a momentum below
    2^-100 or above 2^100) take the full IEEE division and sqrt sequences, the others the
    split forms; the results stay bitwise equal to the oracle either way."""
    nx, ny = 40, 130
    gr, h, hu, hv, B = wet_state(nx, ny, 91 + 3 * recon + riemann, bathy, 2.5, 8.125)
    if case == "dry_cells":
        h[9, 30] = h[20, 95] = h[31, 7] = 0.0   # isolated: ROE never divides 0 by 0
        hu[9, 30] = hv[20, 95] = 0.0
    elif case == "tiny_momentum":
        hu[12, 40:50] = 3e-40
        hv[25, 100:104] = -1e-35
    else:
        hv[15, 60:63] = 4e31
    oracle.mod_apply_bc(gr, h, hu, hv)
    with solver(hip, gr, bathy, recon, riemann) as s:
        s.upload(h, hu, hv, B)
        got = s.rhs()
        rc, want = oracle.mod_rhs(gr, bathy, recon, riemann, h.copy(), hu.copy(), hv.copy(), B)
        assert rc == 0
        for g_, w_, nm in zip(got, want, ("h", "hu", "hv")):
            assert same(g_, w_), (case, "rhs", nm)
        if case == "huge_momentum":
            return  # fluxes of ~1e63 leave no stable step to compare
        dt = 0.02 * min(gr.dx, gr.dy)
        s.upload(h, hu, hv, B)
        ref = [a.copy() for a in (h, hu, hv)]
        rc = oracle.mod_run(gr, bathy, recon, riemann, *ref, B, 2, dt)
        if rc != 0:
            with pytest.raises(hip.SweError):
                s.step(dt, 2)
            return
        s.step(dt, 2)
        for g_, w_, nm in zip(s.download(), ref, ("h", "hu", "hv")):
            assert same(g_, w_), (case, "steps", nm)


@pytest.mark.parametrize("bathy,recon,riemann", SUPPORTED)
def test_reference_fixtures(hip, bathy, recon, riemann):
    """Directly against outputs of the reference's own modular code (wet case)."""
    z = np.load(GOLD)
    nx, ny, lx, ly = z["grid"]
    gr = oracle.grid(int(nx), int(ny), lx=float(lx), ly=float(ly))
    h, hu, hv, B = (z[f"in_{k}_wet"] for k in ("h", "hu", "hv", "B"))
    tag = f"wet_{bathy}_{RECONS[recon]}_{RIEMANNS[riemann]}"
    with solver(hip, gr, bathy, recon, riemann) as s:
        s.upload(h, hu, hv, B)
        assert same(np.stack(s.rhs()), z[f"rhs_{tag}"])
        s.upload(h, hu, hv, B)
        s.step(float(z["stepdt_wet"]), int(z["steps"]))
        assert same(np.stack(s.download()), z[f"run_{tag}"])


@pytest.mark.parametrize("bathy,recon,riemann", [(0, RECON_MUSCL, RIEMANN_ROE),
                                                 (1, RECON_PC, RIEMANN_HLL)])
def test_graph_and_direct_launches_agree(hip, bathy, recon, riemann):
    gr, h, hu, hv, B = wet_state(33, 75, 5, bathy, 2.0625, 4.6875)
    dt = 0.1 * min(gr.dx, gr.dy)
    outs = []
    for graph in (True, False):
        with solver(hip, gr, bathy, recon, riemann) as s:
            s.set_graph(graph)
            s.upload(h, hu, hv, B)
            s.step(dt, 9)
            s.step(dt, 40)   # one replay of the 32-step graph and 8 direct steps
            outs.append(s.download())
    for a, b in zip(*outs):
        assert same(a, b)


@pytest.mark.parametrize("recon", list(RECONS))
@pytest.mark.parametrize("riemann", list(RIEMANNS))
def test_lake_at_rest_plain(hip, recon, riemann):
    nx, ny = 30, 70
    gr = oracle.grid(nx, ny, lx=3.0, ly=7.0)
    h = np.full((nx + 4, ny + 4), 1.25)
    z = np.zeros_like(h)
    with solver(hip, gr, 0, recon, riemann) as s:
        s.upload(h, z, z, z)
        for a in s.rhs():
            assert not a.any()


def test_failures_are_reported(hip):
    """The plain path has no positivity fix: with dry cells HLL finds no dissipation speed
    in the third stage's right-hand side of the first step (the modular oracle's stage-1 and
    stage-2 right-hand sides succeed, the stage-3 one fails), where the reference throws
    (HLL.hpp:23); the HIP step reports it at that stage and step and stops.  (The reference
    throws before it writes U^{n+1}; the fused stage-3 kernel has written it by the time it
    knows, so the arrays left behind differ -- the reference's app never reads them: it ends
    the run, apps/main.cpp:23.)"""
    z = np.load(GOLD)
    nx, ny, lx, ly = z["grid"]
    gr = oracle.grid(int(nx), int(ny), lx=float(lx), ly=float(ly))
    h, hu, hv, B = (z[f"in_{k}_dry"] for k in ("h", "hu", "hv", "B"))
    dt = float(z["stepdt_dry"])
    assert int(z["runrc_dry_0_muscl_hll"]) != 0
    # the failing stage from the oracle's own pieces: rhs of U, U1, U2
    U = [h.copy(), hu.copy(), hv.copy()]
    rc, L = oracle.mod_rhs(gr, 0, RECON_MUSCL, RIEMANN_HLL, *U, B)
    assert rc == 0
    U1 = [1.0 * x + dt * l for x, l in zip(U, L)]
    rc, L = oracle.mod_rhs(gr, 0, RECON_MUSCL, RIEMANN_HLL, *U1, B)
    assert rc == 0
    U2 = [3.0 / 4.0 * x + 1.0 / 4.0 * y + dt / 4.0 * l for x, y, l in zip(U, U1, L)]
    rc, _ = oracle.mod_rhs(gr, 0, RECON_MUSCL, RIEMANN_HLL, *U2, B)
    assert rc != 0
    with solver(hip, gr, 0, RECON_MUSCL, RIEMANN_HLL) as s:
        s.upload(h, hu, hv, B)
        with pytest.raises(hip.SweError) as e:
            s.step(dt, int(z["steps"]))
        assert e.value.code == hip.SWE_ERR_NONFINITE
        assert s.last_status() == (1, 3)


def test_rejected_combinations(hip):
    gr = oracle.grid(8, 8)
    with hip.HipSolver(8, 8, gr.dx, gr.dy, gr.g, hip.BC_REFLECTING) as s:
        with pytest.raises(hip.SweError):
            s.set_scheme(hip.SCHEME_SERIAL, RECON_MUSCL, RIEMANN_ROE)  # bathymetry: HLL only
        with pytest.raises(hip.SweError):
            s.set_scheme(7, RECON_MUSCL, RIEMANN_HLL)
        s.set_scheme(hip.SCHEME_SERIAL_PLAIN, RECON_PC, RIEMANN_ROE)
        assert s.get_scheme() == (hip.SCHEME_SERIAL_PLAIN, RECON_PC, RIEMANN_ROE)
        z = np.ones((12, 12))
        s.upload(z, 0 * z, 0 * z, 0 * z)
        with pytest.raises(hip.SweError):
            s.step_adaptive(0.4, 2)  # the serial family is fixed-dt
        with pytest.raises(hip.SweError):
            hip.HipAdjoint(s, 2, 2)  # the adjoint differentiates the optimized scheme
        s.set_scheme(hip.SCHEME_OPTIMIZED)
        assert s.get_scheme()[0] == hip.SCHEME_OPTIMIZED
    with hip.HipSolver(8, 8, gr.dx, gr.dy, gr.g, hip.BC_TRANSMISSIVE) as s:
        with pytest.raises(hip.SweError):
            s.set_scheme(hip.SCHEME_SERIAL_PLAIN, RECON_PC, RIEMANN_HLL)


@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("bathy,recon,riemann", SUPPORTED)
def test_planted_values_sweep(hip, bathy, recon, riemann, seed):
    """Random shapes around the 60-column strips with the values that decide the split
    forms, the median slopes and the all-wet votes planted: momenta and bottoms of -0/+0,
    flat stretches, equal neighbours; bitwise against the oracle (rhs and 3 steps)."""
    rng = np.random.default_rng(500 + 31 * seed + 7 * recon + riemann + 100 * bathy)
    nx = int(rng.choice([2, 5, 17, 60, 61, 97]))
    ny = int(rng.choice([2, 30, 60, 61, 62, 121]))
    lx, ly = 0.0625 * nx, 0.125 * ny
    gr, h, hu, hv, B = wet_state(nx, ny, 800 + seed, bathy, lx, ly)
    pick = lambda frac: rng.random(h.shape) < frac
    hu[pick(0.05)] = -0.0
    hv[pick(0.05)] = 0.0
    hv[pick(0.05)] = -0.0
    i0, j0 = rng.integers(0, h.shape[0]), rng.integers(0, h.shape[1])
    h[i0:i0 + 3, :] = h[i0, 0]
    hu[:, j0:j0 + 4] = hu[0, j0]
    if bathy:
        B[pick(0.05)] = -0.0
        B[i0:i0 + 5, :] = B[i0, 0]
        oracle.mod_apply_bc(gr, B)
    oracle.mod_apply_bc(gr, h, hu, hv)
    with solver(hip, gr, bathy, recon, riemann, int(rng.choice([1, 4, 32]))) as s:
        s.upload(h, hu, hv, B)
        got = s.rhs()
        rc, want = oracle.mod_rhs(gr, bathy, recon, riemann, h.copy(), hu.copy(), hv.copy(), B)
        assert rc == 0
        for g_, w_, nm in zip(got, want, ("h", "hu", "hv")):
            assert same(g_, w_), (nx, ny, "rhs", nm)
        dt = 0.1 * min(gr.dx, gr.dy)
        s.upload(h, hu, hv, B)
        ref = [a.copy() for a in (h, hu, hv)]
        assert oracle.mod_run(gr, bathy, recon, riemann, *ref, B, 3, dt) == 0
        s.step(dt, 3)
        for g_, w_, nm in zip(s.download(), ref, ("h", "hu", "hv")):
            assert same(g_, w_), (nx, ny, "steps", nm)
