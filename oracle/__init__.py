"""ctypes bindings for the CPU oracle.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg may import this package, and only as the checker (or as the timed
CPU baseline).  The product path (``swe_hip``) never imports it.

Two libraries:

* ``liboracle.so`` -- this build's own C restatement (oracle/swe_oracle.c) of the
  reference's ``include/backend/openmp/omp_kernels.hpp`` plus the ``rk3_step`` driver
  of ``include/backend/openmp/omp_solver.hpp:250-338``.  g is a run-time parameter.
* ``libswe_tangent.so`` -- a forward-mode (dual-number) tangent of the reference's serial
  optimized solver (oracle/swe_tangent.cpp; include/backend/serial/HPC/), the checker of
  the GPU discrete adjoint.  Its primal is bitwise the oracle's fixed-dt run.
* ``_ref/libswe_ref.so`` -- the reference's own header-only kernels compiled where they
  lie (oracle/ref_driver.cpp + ``make -C oracle ref``).  Present only where the
  reference tree was available at build time (the .so itself travels with the tree).
  g is fixed at the reference's 9.80665; ``_ref/libswe_ref_g981.so`` is the same build
  with g = 9.81 (config 1 as stated) through the shadow ``g981/include/constants.hpp``.
* ``_ref/libswe_ref_modular.so`` -- likewise, the reference's modular serial solver parts
  (reconstructions, Riemann solvers, flux assemblies, SSP-RK3; oracle/ref_modular.cpp).

``liboracle.so`` also holds the restatement of the modular serial solver
(oracle/swe_modular.c; the ``mod_*`` wrappers below), the checker of the HIP backend's
modular scheme variants.

Arrays are float64, C-contiguous, shape ``(nx + 2*ng, ny + 2*ng)`` in the reference
layout (row-major, ``k = i*(ny+2*ng) + j``, i = x index, j = y index).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
TANGENT_SO = os.path.join(HERE, "libswe_tangent.so")
REF_SO = os.path.join(HERE, "_ref", "libswe_ref.so")
REF_G981_SO = os.path.join(HERE, "_ref", "libswe_ref_g981.so")
REF_MOD_SO = os.path.join(HERE, "_ref", "libswe_ref_modular.so")

# modular scheme selectors (oracle/swe_modular.c; reconstruction / Riemann solver)
RECON_MUSCL, RECON_PC = 0, 1
RIEMANN_HLL, RIEMANN_RUSANOV, RIEMANN_ROE = 0, 1, 2

G_REF = 9.80665  # include/constants.hpp:4
BC_REFLECTING = 0
BC_TRANSMISSIVE = 1

_dp = C.POINTER(C.c_double)


class OracleGrid(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("ng", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("g", C.c_double),
        ("bc_kind", C.c_int32),
    ]


def build(ref: bool | None = None) -> None:
    """Compile liboracle.so (and the reference harness when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "-B", "liboracle.so", "libswe_tangent.so"],
                   check=True)
    ref_root = os.environ.get("SWE_REFERENCE_ROOT", "/root/reference")
    if ref is None:
        ref = os.path.isdir(os.path.join(ref_root, "include", "backend", "openmp"))
    if ref:
        subprocess.run(["make", "-s", "-C", HERE, "-B", "ref", f"REF={ref_root}"], check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
        L = C.CDLL(ORACLE_SO)
        L.oracle_minmod.restype = C.c_double
        L.oracle_minmod.argtypes = [C.c_double, C.c_double]
        for nm in ("oracle_hll_x", "oracle_hll_y"):
            getattr(L, nm).restype = None
            getattr(L, nm).argtypes = [C.c_double] * 7 + [_dp]
        for nm in ("oracle_x_fluxes", "oracle_y_fluxes"):
            getattr(L, nm).restype = None
            getattr(L, nm).argtypes = [C.POINTER(OracleGrid)] + [_dp] * 10
        L.oracle_rhs.restype = C.c_int
        L.oracle_rhs.argtypes = [C.POINTER(OracleGrid)] + [_dp] * 7
        L.oracle_stable_dt.restype = C.c_double
        L.oracle_stable_dt.argtypes = [C.POINTER(OracleGrid), _dp, _dp, _dp, C.c_double, _dp]
        L.oracle_enforce_positivity.restype = C.c_int
        L.oracle_enforce_positivity.argtypes = [C.POINTER(OracleGrid), _dp, _dp, _dp]
        L.oracle_apply_bc.restype = None
        L.oracle_apply_bc.argtypes = [C.POINTER(OracleGrid), _dp, _dp, _dp]
        L.oracle_run.restype = C.c_int
        L.oracle_mod_apply_bc.restype = None
        L.oracle_mod_apply_bc.argtypes = [C.POINTER(OracleGrid), _dp, _dp, _dp]
        L.oracle_mod_rhs.restype = C.c_int
        L.oracle_mod_rhs.argtypes = [C.POINTER(OracleGrid)] + [C.c_int] * 3 + [_dp] * 7
        L.oracle_mod_run.restype = C.c_int
        L.oracle_mod_run.argtypes = ([C.POINTER(OracleGrid)] + [C.c_int] * 3 + [_dp] * 4 +
                                     [C.c_int64, C.c_double])
        L.oracle_mod_stable_dt.restype = C.c_double
        L.oracle_mod_stable_dt.argtypes = [C.POINTER(OracleGrid), _dp, _dp, _dp, C.c_double]
        L.oracle_mod_riemann.restype = None
        L.oracle_mod_riemann.argtypes = [C.c_double, C.c_int, C.c_int, _dp, _dp, _dp,
                                         C.POINTER(C.c_int)]
        L.oracle_run.argtypes = [C.POINTER(OracleGrid), _dp, _dp, _dp, _dp, C.c_int, C.c_double,
                                 C.c_long, _dp, C.POINTER(C.c_long), C.POINTER(C.c_int)]
        L.oracle_num_threads.restype = C.c_int
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    """The reference harness, or None when it was not built."""
    global _ref
    if _ref is None and ref_available():
        _ref = _load_ref(REF_SO)
    return _ref


_ref_g981 = None


def ref_for_g(g: float):
    """The reference harness compiled with gravity g (9.80665, the reference's own value,
    or 9.81 through the shadow constants header), or None when that build is absent."""
    global _ref_g981
    if g == G_REF:
        return ref()
    if g == 9.81:
        if _ref_g981 is None and os.path.exists(REF_G981_SO):
            _ref_g981 = _load_ref(REF_G981_SO)
        return _ref_g981
    return None


def _load_ref(path):
    R = C.CDLL(path)
    R.ref_g.restype = C.c_double
    R.ref_num_threads.restype = C.c_int
    for nm in ("ref_minmod", "hpc_minmod"):
        getattr(R, nm).restype = C.c_double
        getattr(R, nm).argtypes = [C.c_double, C.c_double]
    for nm in ("ref_hll_x", "ref_hll_y"):
        getattr(R, nm).restype = None
        getattr(R, nm).argtypes = [C.c_double] * 6 + [_dp]
    grid = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]
    for nm in ("ref_faces_and_rhs", "hpc_faces_and_rhs"):
        getattr(R, nm).restype = None
        getattr(R, nm).argtypes = grid + [_dp] * 4 + [C.POINTER(_dp), C.POINTER(_dp)]
    R.ref_stable_dt.restype = C.c_double
    R.ref_stable_dt.argtypes = grid + [_dp] * 3 + [C.c_double]
    R.ref_enforce_positivity.restype = C.c_int
    R.ref_enforce_positivity.argtypes = grid + [_dp] * 3
    R.ref_reflect.restype = None
    R.ref_reflect.argtypes = grid + [_dp] * 3
    R.ref_run.restype = C.c_int
    R.ref_run.argtypes = grid + [_dp] * 4 + [C.c_int, C.c_double, C.c_long, _dp]
    R.ref_initial_state.restype = C.c_int
    R.ref_initial_state.argtypes = ([C.c_int] * 3 + [C.c_double] * 2 + [C.c_int, _dp, C.c_int,
                                                                         _dp] + [_dp] * 4)
    return R


_refmod = None


def ref_modular():
    """The reference modular-solver harness, or None when it was not built."""
    global _refmod
    if _refmod is None and os.path.exists(REF_MOD_SO):
        R = C.CDLL(REF_MOD_SO)
        dims = [C.c_int, C.c_int, C.c_double, C.c_double]
        R.ref_mod_rhs.restype = C.c_int
        R.ref_mod_rhs.argtypes = dims + [C.c_int] * 3 + [_dp] * 7
        R.ref_mod_run.restype = C.c_int
        R.ref_mod_run.argtypes = dims + [C.c_int] * 3 + [_dp] * 4 + [C.c_long, C.c_double]
        R.ref_mod_stable_dt.restype = C.c_double
        R.ref_mod_stable_dt.argtypes = dims + [_dp] * 3 + [C.c_double]
        R.ref_mod_riemann.restype = C.c_int
        R.ref_mod_riemann.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
        _refmod = R
    return _refmod


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def grid(nx, ny, ng=2, dx=None, dy=None, g=G_REF, bc=BC_REFLECTING, lx=1.0, ly=1.0):
    dx = lx / nx if dx is None else dx
    dy = ly / ny if dy is None else dy
    return OracleGrid(nx, ny, ng, dx, dy, g, bc)


# ---------------------------------------------------------------- oracle wrappers
def minmod(a: float, b: float) -> float:
    return lib().oracle_minmod(a, b)


def hll(direction: str, hL, huL, hvL, hR, huR, hvR, g=G_REF):
    out = (C.c_double * 3)()
    fn = lib().oracle_hll_x if direction == "x" else lib().oracle_hll_y
    fn(g, hL, huL, hvL, hR, huR, hvR, out)
    return tuple(out)


def apply_bc(gr: OracleGrid, h, hu=None, hv=None):
    lib().oracle_apply_bc(C.byref(gr), _p(h), _p(hu) if hu is not None else None,
                          _p(hv) if hv is not None else None)


def faces(gr: OracleGrid, h, hu, hv, B):
    """The 12 face arrays fxm(h,hu,hv) fxp(..) fym(..) fyp(..) of the current arrays (no
    ghost fill)."""
    f = [np.zeros_like(h) for _ in range(12)]
    lib().oracle_x_fluxes(C.byref(gr), _p(h), _p(hu), _p(hv), _p(B), *(_p(a) for a in f[:6]))
    lib().oracle_y_fluxes(C.byref(gr), _p(h), _p(hu), _p(hv), _p(B), *(_p(a) for a in f[6:]))
    return f


def rhs(gr: OracleGrid, h, hu, hv, B):
    """Ghost fill (in place) then compute_rhs; returns (rhs_h, rhs_hu, rhs_hv)."""
    out = [np.zeros_like(h) for _ in range(3)]
    rc = lib().oracle_rhs(C.byref(gr), _p(h), _p(hu), _p(hv), _p(B), *(_p(o) for o in out))
    if rc != 0:
        raise MemoryError("oracle_rhs allocation failed")
    return out


def stable_dt(gr: OracleGrid, h, hu, hv, cfl):
    ms = C.c_double()
    dt = lib().oracle_stable_dt(C.byref(gr), _p(h), _p(hu), _p(hv), cfl, C.byref(ms))
    return dt, ms.value


def enforce_positivity(gr: OracleGrid, h, hu, hv) -> int:
    return lib().oracle_enforce_positivity(C.byref(gr), _p(h), _p(hu), _p(hv))


def run_status(gr: OracleGrid, h, hu, hv, B, nsteps, dt=None, cfl=None):
    """Advance in place.  Exactly one of dt (fixed) or cfl (adaptive per step).

    Returns (rc, steps_done, dt_hist, failed_stage): on a non-finite state (rc 1) the run
    stopped in stage failed_stage (1..3) of step steps_done + 1, and (h, hu, hv) are left as
    the reference's throw leaves them (omp_solver.hpp:250-338)."""
    assert (dt is None) != (cfl is None)
    hist = np.zeros(max(nsteps, 1))
    done = C.c_long()
    stage = C.c_int()
    mode, val = (0, dt) if dt is not None else (1, cfl)
    rc = lib().oracle_run(C.byref(gr), _p(h), _p(hu), _p(hv), _p(B), mode, val, nsteps,
                          _p(hist), C.byref(done), C.byref(stage))
    if rc < 0:
        raise MemoryError("oracle_run allocation failed")
    return rc, done.value, hist[:nsteps], stage.value


def run(gr: OracleGrid, h, hu, hv, B, nsteps, dt=None, cfl=None):
    """Advance in place.  Exactly one of dt (fixed) or cfl (adaptive per step).

    Returns (rc, steps_done, dt_hist)."""
    rc, done, hist, _ = run_status(gr, h, hu, hv, B, nsteps, dt=dt, cfl=cfl)
    return rc, done, hist


# ---------------------------------------------------------------- modular serial solver
def mod_apply_bc(gr: OracleGrid, h, hu=None, hv=None):
    """reflecting_walls.hpp ghost fill (y walls first); hu = hv = None: scalar fill."""
    lib().oracle_mod_apply_bc(C.byref(gr), _p(h), _p(hu) if hu is not None else None,
                              _p(hv) if hv is not None else None)


def mod_rhs(gr: OracleGrid, bathy, recon, riemann, h, hu, hv, B):
    """(rc, (rh, rhu, rhv)); fills the ghosts of h, hu, hv like the reference's rhs."""
    r = [np.zeros_like(h) for _ in range(3)]
    Bz = B if B is not None else np.zeros_like(h)
    rc = lib().oracle_mod_rhs(C.byref(gr), int(bathy), int(recon), int(riemann), _p(h), _p(hu),
                              _p(hv), _p(Bz), *(_p(a) for a in r))
    return rc, r


def mod_run(gr: OracleGrid, bathy, recon, riemann, h, hu, hv, B, nsteps, dt):
    """Advance in place; 0, 1 (non-finite) or 2 (HLL without dissipation speed)."""
    Bz = B if B is not None else np.zeros_like(h)
    return lib().oracle_mod_run(C.byref(gr), int(bathy), int(recon), int(riemann), _p(h),
                                _p(hu), _p(hv), _p(Bz), int(nsteps), float(dt))


def mod_stable_dt(gr: OracleGrid, h, hu, hv, cfl):
    return lib().oracle_mod_stable_dt(C.byref(gr), _p(h), _p(hu), _p(hv), cfl)


def mod_riemann(direction: str, kind, L, R, g=G_REF):
    F = np.zeros(3)
    fail = C.c_int()
    lib().oracle_mod_riemann(g, int(direction == "y"), int(kind), _p(np.asarray(L, float)),
                             _p(np.asarray(R, float)), _p(F), C.byref(fail))
    return F, fail.value


def ref_mod_rhs(nx, ny, lx, ly, bathy, recon, riemann, h, hu, hv, B):
    r = [np.zeros_like(h) for _ in range(3)]
    Bz = B if B is not None else np.zeros_like(h)
    rc = ref_modular().ref_mod_rhs(nx, ny, lx, ly, int(bathy), int(recon), int(riemann), _p(h),
                                   _p(hu), _p(hv), _p(Bz), *(_p(a) for a in r))
    return rc, r


def ref_mod_run(nx, ny, lx, ly, bathy, recon, riemann, h, hu, hv, B, nsteps, dt):
    Bz = B if B is not None else np.zeros_like(h)
    return ref_modular().ref_mod_run(nx, ny, lx, ly, int(bathy), int(recon), int(riemann), _p(h),
                                     _p(hu), _p(hv), _p(Bz), int(nsteps), float(dt))


def ref_mod_stable_dt(nx, ny, lx, ly, h, hu, hv, cfl):
    return ref_modular().ref_mod_stable_dt(nx, ny, lx, ly, _p(h), _p(hu), _p(hv), cfl)


def ref_mod_riemann(direction: str, kind, L, R):
    F = np.zeros(3)
    rc = ref_modular().ref_mod_riemann(int(direction == "y"), int(kind),
                                       _p(np.asarray(L, float)), _p(np.asarray(R, float)), _p(F))
    return F, rc


# ---------------------------------------------------------------- reference harness
def ref_faces_and_rhs(gr: OracleGrid, h, hu, hv, B, hpc=False):
    f = [np.zeros_like(h) for _ in range(12)]
    r = [np.zeros_like(h) for _ in range(3)]
    fp = (_dp * 12)(*[_p(a) for a in f])
    rp = (_dp * 3)(*[_p(a) for a in r])
    fn = ref().hpc_faces_and_rhs if hpc else ref().ref_faces_and_rhs
    fn(gr.nx, gr.ny, gr.ng, gr.dx, gr.dy, _p(h), _p(hu), _p(hv), _p(B), fp, rp)
    return f, r


def ref_run(gr: OracleGrid, h, hu, hv, B, nsteps, dt=None, cfl=None, lib=None):
    """The reference kernels' rk3_step loop (lib: a harness from ref_for_g; default the
    reference's own g)."""
    assert (dt is None) != (cfl is None)
    R = lib if lib is not None else ref()
    hist = np.zeros(max(nsteps, 1))
    mode, val = (0, dt) if dt is not None else (1, cfl)
    rc = R.ref_run(gr.nx, gr.ny, gr.ng, gr.dx, gr.dy, _p(h), _p(hu), _p(hv), _p(B), mode, val,
                   nsteps, _p(hist))
    return rc, hist[:nsteps]


# reference enum values (configs/config.hpp:11-13)
IC_TYPES = {"GaussInitial": 0, "StillWater": 1, "DamBreak": 2, "DamBreakRadial": 3}
BATHY_TYPES = {"None": 0, "Flat": 1, "GaussHill": 2}
IC_KEYS = ("peak_height", "x0", "y0", "sigma_x", "sigma_y", "h0", "dam_height", "dam_x",
           "dam_x0", "dam_y0", "dam_radius")
BATHY_KEYS = ("b0", "bathy_x0", "bathy_y0", "bathy_sigma_x", "bathy_sigma_y",
              "bathy_peak_height")


def ref_initial_state(nx, ny, ng, lx, ly, ic_type: str, ic: dict, bathy_type: str, bt: dict):
    """The reference's own setup code (apply_initial_condition / apply_bathymetry) and the
    solver constructor's ghost fills (omp_solver.hpp:118-127): full arrays h, hu, hv, B."""
    shape = (nx + 2 * ng, ny + 2 * ng)
    out = [np.zeros(shape) for _ in range(4)]
    icv = np.array([float(ic.get(k, 0.0)) for k in IC_KEYS])
    btv = np.array([float(bt.get(k, 0.0)) for k in BATHY_KEYS])
    rc = ref().ref_initial_state(nx, ny, ng, lx, ly, IC_TYPES[ic_type], _p(icv),
                                 BATHY_TYPES[bathy_type], _p(btv), *(_p(a) for a in out))
    if rc != 0:
        raise ValueError("reference setup code threw")
    return out


_tan = None


def tangent_lib():
    global _tan
    if _tan is None:
        if not os.path.exists(TANGENT_SO):
            subprocess.run(["make", "-s", "-C", HERE, "libswe_tangent.so"], check=True)
        L = C.CDLL(TANGENT_SO)
        L.tangent_primal.restype = C.c_int
        L.tangent_primal.argtypes = ([C.c_int] * 3 + [C.c_double] * 3 + [C.c_int] + [_dp] * 5 +
                                     [C.c_long])
        L.tangent_run.restype = C.c_int
        L.tangent_run.argtypes = ([C.c_int] * 3 + [C.c_double] * 3 + [C.c_int] + [_dp] * 9 +
                                  [C.c_long])
        _tan = L
    return _tan


def tangent_primal(gr, h, hu, hv, B, dts):
    """Fixed-dt-sequence run of the tangent's primal (T = double); arrays in place."""
    dts = np.ascontiguousarray(dts, dtype=np.float64)
    return tangent_lib().tangent_primal(gr.nx, gr.ny, gr.ng, gr.dx, gr.dy, gr.g, gr.bc_kind,
                                        _p(h), _p(hu), _p(hv), _p(B), _p(dts), len(dts))


def tangent_run(gr, h, hu, hv, B, dh, dhu, dhv, dB, dts):
    """Directional derivative of the final state along (dh, dhu, dhv, dB), dt frozen.
    h/hu/hv become the final state, dh/dhu/dhv its tangent.  Returns 1 on non-finite."""
    dts = np.ascontiguousarray(dts, dtype=np.float64)
    return tangent_lib().tangent_run(gr.nx, gr.ny, gr.ng, gr.dx, gr.dy, gr.g, gr.bc_kind,
                                     _p(h), _p(hu), _p(hv), _p(B), _p(dh), _p(dhu), _p(dhv),
                                     _p(dB), _p(dts), len(dts))
