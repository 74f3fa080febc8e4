"""Seeded synthetic workloads for the shallow-water step (host side, numpy).

Every array is float64 in the reference layout: shape ``(nx + 2*ng, ny + 2*ng)``,
row-major, ``k = i*(ny + 2*ng) + j`` with i the x index and j the y index
(include/backend/openmp/omp_kernels.hpp:33-35; include/core/array2D.hpp:29-34).
Cell centres are ``x = (i + 0.5)*dx`` for interior index i (include/core/grid.hpp:22).

Randomness comes from a counter-based splitmix64 hash so the same seed gives the same
bits on every machine (numpy's Generator streams are not used).  Gaussians use
``np.exp``; tests that pin results by hash also pin the generated inputs by hash.

The workloads:
  dam_break      -- 8192^2 on [0,20]^2, flat bottom, reflecting walls, h = 2 for x < 10
                    else 1, plus 1e-3*U(0,1) noise; CFL 0.45.
  gauss_pulse    -- 256^2 on [0,1]^2, flat bottom, reflecting walls,
                    h = 1 + 0.5*exp(-((x-.5)^2 + (y-.5)^2)/0.01); CFL 0.4, g = 9.81.
  lake_wet_dry   -- 4096^2 on [0,10]^2, transmissive walls, B = 32 seeded Gaussians,
                    lake at rest eta = 0.5 (h = max(0, eta - B)), optional 0.05 bump;
                    eta0 = 2.0 gives the fully wet variant.
  ensemble       -- 512 instances, sides log-uniform in [64, 1024], 8-bump bathymetry
                    and a Gaussian surface each.
  adjoint_case   -- 1024^2, 16-Gaussian bathymetry, fully wet surface with a bump.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """n outputs of splitmix64 at counters 1..n for (seed, stream), as uint64."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * 0x100000001B3 + stream * 0xD1B54A32D192ED03) & 0xFFFFFFFFFFFFFFFF)
        z = base + (np.arange(1, n + 1, dtype=np.uint64) * _GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, n: int, stream: int = 0, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
    """n doubles in [lo, hi) from the top 53 bits of splitmix64."""
    u = (splitmix64(seed, n, stream) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return lo + (hi - lo) * u


def centres(n: int, L: float) -> np.ndarray:
    d = L / n
    return (np.arange(n, dtype=np.float64) + 0.5) * d


def empty_state(nx: int, ny: int, ng: int = 2):
    shape = (nx + 2 * ng, ny + 2 * ng)
    return [np.zeros(shape) for _ in range(4)]  # h, hu, hv, B


def fill_ghosts(a: np.ndarray, ng: int, bc: int = 0, sign_x: float = 1.0, sign_y: float = 1.0):
    """Ghost fill in place, same two-pass order as include/backend/openmp/omp_solver.hpp:340-393.

    bc 0 = reflecting (mirror; multiply by sign_x on x-ghost rows, sign_y on y-ghost
    columns), bc 1 = transmissive (copy the nearest interior cell)."""
    nxt, nyt = a.shape
    nx, ny = nxt - 2 * ng, nyt - 2 * ng
    for g in range(ng):
        iL, iR = ng - 1 - g, ng + nx + g
        sL, sR = (ng + g, ng + nx - 1 - g) if bc == 0 else (ng, ng + nx - 1)
        flip = bc == 0 and sign_x == -1.0
        a[iL, ng:ng + ny] = -a[sL, ng:ng + ny] if flip else a[sL, ng:ng + ny]
        a[iR, ng:ng + ny] = -a[sR, ng:ng + ny] if flip else a[sR, ng:ng + ny]
    for g in range(ng):
        jB, jT = ng - 1 - g, ng + ny + g
        sB, sT = (ng + g, ng + ny - 1 - g) if bc == 0 else (ng, ng + ny - 1)
        flip = bc == 0 and sign_y == -1.0
        a[:, jB] = -a[:, sB] if flip else a[:, sB]
        a[:, jT] = -a[:, sT] if flip else a[:, sT]
    return a


def gaussian_sum(x, y, amps, widths, cx, cy):
    """sum_k A_k exp(-((x-cx_k)^2 + (y-cy_k)^2) / (2 w_k^2)) on the x (rows) by y (cols) grid."""
    out = np.zeros((x.size, y.size))
    for A, w, x0, y0 in zip(amps, widths, cx, cy):
        gx = np.exp(-((x - x0) ** 2) / (2.0 * w * w))
        gy = np.exp(-((y - y0) ** 2) / (2.0 * w * w))
        out += A * np.outer(gx, gy)
    return out


class Case(dict):
    """A single-grid workload: keys nx, ny, ng, lx, ly, dx, dy, g, bc, cfl, steps, h, hu, hv, B."""


def _case(nx, ny, lx, ly, g, bc, cfl, steps, ng=2, name=""):
    h, hu, hv, B = empty_state(nx, ny, ng)
    return Case(name=name, nx=nx, ny=ny, ng=ng, lx=lx, ly=ly, dx=lx / nx, dy=ly / ny, g=g, bc=bc,
                cfl=cfl, steps=steps, h=h, hu=hu, hv=hv, B=B)


def dam_break(n: int = 8192, seed: int = 3, steps: int = 1000, noise: float = 1e-3,
              g: float = 9.80665) -> Case:
    c = _case(n, n, 20.0, 20.0, g, 0, 0.45, steps, name=f"dam_break_{n}")
    ng = c["ng"]
    x = centres(n, 20.0)
    hin = np.where(x < 10.0, 2.0, 1.0)[:, None] * np.ones((1, n))
    if noise:
        hin = hin + noise * uniform(seed, n * n, stream=1).reshape(n, n)
    c["h"][ng:ng + n, ng:ng + n] = hin
    fill_ghosts(c["h"], ng, 0)
    return c


def gauss_pulse(n: int = 256, steps: int = 1000, g: float = 9.81) -> Case:
    c = _case(n, n, 1.0, 1.0, g, 0, 0.4, steps, name=f"gauss_pulse_{n}")
    ng = c["ng"]
    x = centres(n, 1.0)
    r2 = (x[:, None] - 0.5) ** 2 + (x[None, :] - 0.5) ** 2
    c["h"][ng:ng + n, ng:ng + n] = 1.0 + 0.5 * np.exp(-r2 / 0.01)
    fill_ghosts(c["h"], ng, 0)
    return c


def lake_wet_dry(n: int = 4096, seed: int = 2, bump: bool = True, steps: int = 2000,
                 g: float = 9.80665, nbumps: int = 32, L: float = 10.0,
                 eta0: float = 0.5) -> Case:
    """Lake at rest over seeded Gaussian bathymetry.  eta0 = 0.5 leaves part of the domain
    dry (B reaches ~1.6 for the default seed); eta0 = 2.0 keeps every cell wet."""
    c = _case(n, n, L, L, g, 1, 0.4, steps, name=f"lake_wet_dry_{n}")
    ng = c["ng"]
    x = centres(n, L)
    A = uniform(seed, nbumps, 11, 0.1, 0.9)
    w = uniform(seed, nbumps, 12, 0.2, 1.0)
    cx = uniform(seed, nbumps, 13, 0.0, L)
    cy = uniform(seed, nbumps, 14, 0.0, L)
    Bi = gaussian_sum(x, x, A, w, cx, cy)
    eta = np.full((n, n), eta0)
    if bump:
        bx, by = uniform(seed, 2, 15, 0.0, L)
        eta = eta + 0.05 * np.outer(np.exp(-((x - bx) ** 2) / (2 * 0.25 ** 2)),
                                    np.exp(-((x - by) ** 2) / (2 * 0.25 ** 2)))
    c["B"][ng:ng + n, ng:ng + n] = Bi
    c["h"][ng:ng + n, ng:ng + n] = np.maximum(0.0, eta - Bi)
    fill_ghosts(c["B"], ng, 1)
    fill_ghosts(c["h"], ng, 1)
    return c


def adjoint_case(n: int = 1024, seed: int = 5, steps: int = 2000, g: float = 9.80665,
                 L: float = 10.0, nbumps: int = 16, background: bool = True) -> Case:
    """Adjoint workload: 16 seeded Gaussian bumps of bathymetry (A in [0.05, 0.3], width in
    [0.5, 1.5]), surface eta = 1 + 0.1 * (Gaussian bump of width 0.5 at a seeded place),
    reflecting walls.

    background=True adds a smooth surface undulation (0.01 sin*cos + a 0.003 ramp) and a
    background flow (u = 0.1 sin + 0.02, v = 0.1 sin - 0.03).  Without it, most of the lake
    is exactly at rest with eta = 1.0 to the last bit: neighbouring differences are exactly
    0, where minmod's directional derivative minmod(da, db) is not linear in the direction,
    so the discrete map has no derivative there (and the branch-taken linearization of the
    rounding-level staircase at the bump's edge grows without bound over 2000 steps).
    """
    c = _case(n, n, L, L, g, 0, 0.4, steps, name=f"adjoint_{n}")
    ng = c["ng"]
    x = centres(n, L)
    A = uniform(seed, nbumps, 21, 0.05, 0.3)
    w = uniform(seed, nbumps, 22, 0.5, 1.5)
    cx = uniform(seed, nbumps, 23, 0.0, L)
    cy = uniform(seed, nbumps, 24, 0.0, L)
    Bi = gaussian_sum(x, x, A, w, cx, cy)
    bx, by = uniform(seed, 2, 25, 0.3 * L, 0.7 * L)
    eta = 1.0 + 0.1 * np.outer(np.exp(-((x - bx) ** 2) / (2 * 0.5 ** 2)),
                               np.exp(-((x - by) ** 2) / (2 * 0.5 ** 2)))
    if background:
        eta = eta + 0.01 * np.outer(np.sin(2 * np.pi * x / L + 0.3),
                                    np.cos(2 * np.pi * x / L + 0.7)) + 0.003 * (x[:, None] / L)
    c["B"][ng:ng + n, ng:ng + n] = Bi
    h = np.maximum(0.0, eta - Bi)
    c["h"][ng:ng + n, ng:ng + n] = h
    if background:
        u = 0.1 * np.outer(np.ones(n), np.sin(2 * np.pi * x / L + 1.1)) + 0.02
        v = 0.1 * np.outer(np.sin(2 * np.pi * x / L + 0.4), np.ones(n)) - 0.03
        c["hu"][ng:ng + n, ng:ng + n] = h * u
        c["hv"][ng:ng + n, ng:ng + n] = h * v
    fill_ghosts(c["B"], ng, 0)
    fill_ghosts(c["h"], ng, 0)
    fill_ghosts(c["hu"], ng, 0, sign_x=-1.0)
    fill_ghosts(c["hv"], ng, 0, sign_y=-1.0)
    return c


def adjoint_gauges(n: int, seed: int = 5, count: int = 16):
    """Seeded gauge cells (row, column index arrays in the padded layout, ng = 2)."""
    r = uniform(seed, 2 * count, 26, 0.0, 1.0)
    rows = 2 + np.minimum((r[:count] * n).astype(np.int64), n - 1)
    cols = 2 + np.minimum((r[count:] * n).astype(np.int64), n - 1)
    return rows, cols


def ensemble_sides(count: int = 512, seed: int = 4, lo: int = 64, hi: int = 1024) -> np.ndarray:
    """Square sides log-uniform in [lo, hi]."""
    u = uniform(seed, count, 31)
    return np.rint(lo * (hi / lo) ** u).astype(np.int64)


def ensemble_member(idx: int, n: int, seed: int = 4, g: float = 9.80665, steps: int = 500) -> Case:
    """Instance idx of the ensemble: [0,1]^2, reflecting, 8-bump bathymetry (A in 0.05..0.3),
    Gaussian surface bump on eta = 1 (fully wet)."""
    c = _case(n, n, 1.0, 1.0, g, 0, 0.4, steps, name=f"member_{idx}_{n}")
    ng = c["ng"]
    x = centres(n, 1.0)
    s = 1000 + idx
    A = uniform(seed, 8, s * 8 + 1, 0.05, 0.3)
    w = uniform(seed, 8, s * 8 + 2, 0.05, 0.2)
    cx = uniform(seed, 8, s * 8 + 3, 0.0, 1.0)
    cy = uniform(seed, 8, s * 8 + 4, 0.0, 1.0)
    Bi = gaussian_sum(x, x, A, w, cx, cy)
    px, py = uniform(seed, 2, s * 8 + 5, 0.25, 0.75)
    amp = uniform(seed, 1, s * 8 + 6, 0.05, 0.2)[0]
    eta = 1.0 + amp * np.outer(np.exp(-((x - px) ** 2) / (2 * 0.08 ** 2)),
                               np.exp(-((x - py) ** 2) / (2 * 0.08 ** 2)))
    c["B"][ng:ng + n, ng:ng + n] = Bi
    c["h"][ng:ng + n, ng:ng + n] = np.maximum(0.0, eta - Bi)
    fill_ghosts(c["B"], ng, 0)
    fill_ghosts(c["h"], ng, 0)
    return c


def random_state(nx: int, ny: int, seed: int, ng: int = 2, wet_dry: bool = False,
                 bathy: bool = True, g: float = 9.80665, lx: float = 1.0, ly: float = 1.0,
                 bc: int = 0) -> Case:
    """Rough random state for kernel-level parity (all HLL branches, slope sign changes)."""
    c = _case(nx, ny, lx, ly, g, bc, 0.4, 1, ng, name=f"random_{nx}x{ny}_{seed}")
    nxt, nyt = nx + 2 * ng, ny + 2 * ng
    n = nxt * nyt
    u = lambda s, lo, hi: uniform(seed, n, s, lo, hi).reshape(nxt, nyt)
    B = u(41, 0.0, 0.6) if bathy else np.zeros((nxt, nyt))
    h = u(42, 0.2, 1.5)
    if wet_dry:
        dry = u(43, 0.0, 1.0) < 0.3
        h = np.where(dry, u(44, 0.0, 2e-8), h)
    hu = u(45, -1.0, 1.0) * h * 2.0
    hv = u(46, -1.0, 1.0) * h * 2.0
    # a few supercritical patches so both one-sided HLL branches appear
    fast = u(47, 0.0, 1.0) < 0.1
    hu = np.where(fast, h * 8.0 * np.sign(u(48, -1.0, 1.0)), hu)
    c.update(h=np.ascontiguousarray(h), hu=np.ascontiguousarray(hu),
             hv=np.ascontiguousarray(hv), B=np.ascontiguousarray(B))
    return c
