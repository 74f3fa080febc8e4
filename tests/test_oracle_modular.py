"""The modular serial solver restatement (oracle/swe_modular.c) is pinned (CPU only).

* bitwise against the committed fixtures tests/golden/modular.npz, written from the
  reference's own modular headers (tests/golden/make_golden_modular.py);
* bitwise against those headers compiled here (oracle/_ref/libswe_ref_modular.so) on fresh
  random inputs, when the harness was built;
* closed forms: lake at rest, consistency of every Riemann solver, zero mass flux through
  reflecting walls (mass telescoping), the CFL-check formula at rest.
"""
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)

import oracle  # noqa: E402
from oracle import RECON_MUSCL, RECON_PC, RIEMANN_HLL, RIEMANN_ROE, RIEMANN_RUSANOV  # noqa: E402

GOLD = os.path.join(REPO, "tests", "golden", "modular.npz")
RECONS = {RECON_MUSCL: "muscl", RECON_PC: "pc"}
RIEMANNS = {RIEMANN_HLL: "hll", RIEMANN_RUSANOV: "rusanov", RIEMANN_ROE: "roe"}
COMBOS = [(b, r, q) for b in (0, 1) for r in RECONS for q in RIEMANNS]


def same(a, b):
    """Bitwise equal, except that NaN matches NaN (the payload and sign of a NaN depend on
    which operation produced it first, which differs between compilations)."""
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    if a.shape != b.shape:
        return False
    na, nb = np.isnan(a), np.isnan(b)
    return bool(np.array_equal(na, nb) and
                np.array_equal(a[~na].view(np.uint64), b[~nb].view(np.uint64)))


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


@pytest.mark.parametrize("direction", ["x", "y"])
@pytest.mark.parametrize("kind", list(RIEMANNS))
def test_riemann_scalar_golden(gold, direction, kind):
    L, R = gold["riemann_L"], gold["riemann_R"]
    want = gold[f"riemann_{direction}_{RIEMANNS[kind]}"]
    want_fail = gold[f"riemann_{direction}_{RIEMANNS[kind]}_fail"]
    for i in range(len(L)):
        F, fail = oracle.mod_riemann(direction, kind, L[i], R[i])
        assert fail == want_fail[i], i
        if not fail:
            assert same(F, want[i]), (i, F, want[i])


@pytest.mark.parametrize("case", ["wet", "dry"])
@pytest.mark.parametrize("bathy,recon,riemann", COMBOS)
def test_rhs_and_steps_golden(gold, case, bathy, recon, riemann):
    nx, ny, lx, ly = gold["grid"]
    gr = oracle.grid(int(nx), int(ny), lx=float(lx), ly=float(ly))
    ins = [gold[f"in_{k}_{case}"].copy() for k in ("h", "hu", "hv", "B")]
    tag = f"{case}_{bathy}_{RECONS[recon]}_{RIEMANNS[riemann]}"
    s = [a.copy() for a in ins[:3]]
    rc, r = oracle.mod_rhs(gr, bathy, recon, riemann, *s, ins[3])
    assert (rc != 0) == (int(gold[f"rhsrc_{tag}"]) != 0)
    if rc == 0:
        assert same(np.stack(r), gold[f"rhs_{tag}"])
    s = [a.copy() for a in ins[:3]]
    rc = oracle.mod_run(gr, bathy, recon, riemann, *s, ins[3], int(gold["steps"]),
                        float(gold[f"stepdt_{case}"]))
    assert (rc != 0) == (int(gold[f"runrc_{tag}"]) != 0)
    if rc == 0:
        assert same(np.stack(s), gold[f"run_{tag}"])


@pytest.mark.parametrize("case", ["wet", "dry"])
def test_cfl_formula_golden(gold, case):
    nx, ny, lx, ly = gold["grid"]
    gr = oracle.grid(int(nx), int(ny), lx=float(lx), ly=float(ly))
    h, hu, hv = (gold[f"in_{k}_{case}"].copy() for k in ("h", "hu", "hv"))
    assert same(oracle.mod_stable_dt(gr, h, hu, hv, 0.4), gold[f"dt_{case}"])


@pytest.mark.skipif(oracle.ref_modular() is None, reason="reference harness not built")
@pytest.mark.parametrize("bathy,recon,riemann", COMBOS)
def test_against_reference_headers(bathy, recon, riemann):
    rng = np.random.default_rng(100 + 10 * bathy + 3 * recon + riemann)
    for nx, ny in ((9, 7), (23, 31), (40, 12)):
        lx, ly = 0.125 * nx, 0.25 * ny
        gr = oracle.grid(nx, ny, lx=lx, ly=ly)
        shape = (nx + 4, ny + 4)
        h = 0.3 + rng.random(shape)
        dry = rng.random(shape) < 0.1
        h[dry] = 1e-8 * rng.integers(0, 3, int(dry.sum()))   # 0, eps or 2 eps
        hu, hv = 0.4 * rng.standard_normal(shape), 0.4 * rng.standard_normal(shape)
        B = 0.3 * rng.random(shape) if bathy else np.zeros(shape)
        oracle.mod_apply_bc(gr, B)
        a, b = [x.copy() for x in (h, hu, hv)], [x.copy() for x in (h, hu, hv)]
        rc1, r1 = oracle.mod_rhs(gr, bathy, recon, riemann, *a, B)
        rc2, r2 = oracle.ref_mod_rhs(nx, ny, lx, ly, bathy, recon, riemann, *b, B)
        assert (rc1 != 0) == (rc2 != 0)
        assert all(same(x, y) for x, y in zip(a, b))  # ghost fill
        if rc1 == 0:
            assert all(same(x, y) for x, y in zip(r1, r2))
        a, b = [x.copy() for x in (h, hu, hv)], [x.copy() for x in (h, hu, hv)]
        dt = 0.02 * min(lx / nx, ly / ny)
        q1 = oracle.mod_run(gr, bathy, recon, riemann, *a, B, 4, dt)
        q2 = oracle.ref_mod_run(nx, ny, lx, ly, bathy, recon, riemann, *b, B, 4, dt)
        assert (q1 != 0) == (q2 != 0)
        if q1 == 0:
            assert all(same(x, y) for x, y in zip(a, b))


# ------------------------------------------------------------------ closed forms
@pytest.mark.parametrize("bathy,recon,riemann", COMBOS)
def test_lake_at_rest(bathy, recon, riemann):
    """Still water: flat without bathymetry gives rhs == 0 exactly; with bathymetry the
    hydrostatic reconstruction keeps it to round-off (HLL's star flux rounds)."""
    nx, ny = 20, 16
    gr = oracle.grid(nx, ny, lx=2.5, ly=2.0)
    rng = np.random.default_rng(7)
    shape = (nx + 4, ny + 4)
    B = 0.5 * rng.random(shape) if bathy else np.zeros(shape)
    oracle.mod_apply_bc(gr, B)
    h = 1.5 - B
    z = np.zeros(shape)
    rc, r = oracle.mod_rhs(gr, bathy, recon, riemann, h, z.copy(), z.copy(), B)
    assert rc == 0
    tol = 0.0 if not bathy else 1e-12
    for a in r:
        assert np.abs(a).max() <= tol


@pytest.mark.parametrize("direction", ["x", "y"])
@pytest.mark.parametrize("kind", list(RIEMANNS))
def test_riemann_consistency(direction, kind):
    """F(U, U) equals the physical flux: exactly for Rusanov and ROE (the dissipation
    vanishes), to a few ulps for HLL's star branch."""
    g = oracle.G_REF
    rng = np.random.default_rng(11)
    for _ in range(200):
        U = np.array([0.2 + rng.random(), rng.normal(0, 0.5), rng.normal(0, 0.5)])
        F, fail = oracle.mod_riemann(direction, kind, U, U)
        assert not fail
        u, v = U[1] / U[0], U[2] / U[0]
        if direction == "x":
            phys = np.array([U[1], u * u * U[0] + 0.5 * g * U[0] * U[0], U[1] * v])
        else:
            phys = np.array([U[2], U[1] * v, U[0] * v * v + 0.5 * g * U[0] * U[0]])
        if kind == RIEMANN_HLL:
            np.testing.assert_allclose(F, phys, rtol=1e-14, atol=1e-15)
        else:
            assert same(F, phys)


@pytest.mark.parametrize("bathy,recon,riemann", COMBOS)
def test_mass_telescopes(bathy, recon, riemann):
    """Reflecting walls carry no mass flux, so the interior sum of the mass rhs is 0 up to
    the rounding of the sum."""
    nx, ny = 24, 18
    gr = oracle.grid(nx, ny, lx=3.0, ly=2.25)
    rng = np.random.default_rng(3)
    shape = (nx + 4, ny + 4)
    h = 0.8 + 0.4 * rng.random(shape)
    hu, hv = 0.2 * rng.standard_normal(shape), 0.2 * rng.standard_normal(shape)
    B = 0.2 * rng.random(shape) if bathy else np.zeros(shape)
    oracle.mod_apply_bc(gr, B)
    rc, r = oracle.mod_rhs(gr, bathy, recon, riemann, h, hu, hv, B)
    assert rc == 0
    rh = r[0][2:-2, 2:-2]
    assert abs(rh.sum()) <= 1e-12 * np.abs(rh).sum()


def test_cfl_formula_at_rest():
    nx, ny = 10, 6
    gr = oracle.grid(nx, ny, lx=1.0, ly=0.75)
    h = np.full((nx + 4, ny + 4), 2.0)
    z = np.zeros_like(h)
    c = np.sqrt(oracle.G_REF * 2.0)
    want = 0.4 / ((0.0 + c) / gr.dx + (0.0 + c) / gr.dy)
    assert oracle.mod_stable_dt(gr, h, z, z, 0.4) == want
