"""Python binding of libswe_hip.so (the MI355X shallow-water backend) over its C ABI.

The C ABI is declared in include/swe_hip.h.  This module only marshals numpy arrays
(host) or torch tensors (device pointers) through ctypes; all compute runs in the HIP
kernels.  There is no CPU fallback: if the library is missing or no GPU is visible the
calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# SWE_HIP_LIB overrides the library path (used to time alternative builds side by side).
LIB_PATH = os.environ.get("SWE_HIP_LIB") or os.path.join(HERE, "libswe_hip.so")

SWE_OK = 0
SWE_ERR_INVALID = 1
SWE_ERR_HIP = 2
SWE_ERR_NONFINITE = 3
SWE_ERR_CFL = 4
SWE_ERR_OOM = 5
SWE_ERR_NO_DEVICE = 6

BC_REFLECTING = 0
BC_TRANSMISSIVE = 1

PRECISION_EXACT = 0   # bitwise the reference's arithmetic
PRECISION_FAST = 1    # reciprocal-based velocities, few-ulp differences per step

SCHEME_OPTIMIZED = 0     # MUSCL + HLL + hydrostatic reconstruction (omp_kernels.hpp)
SCHEME_SERIAL = 1        # modular SerialSolver with bathymetry (HLL; MUSCL or PC)
SCHEME_SERIAL_PLAIN = 2  # modular SerialSolver without bathymetry (any recon x Riemann)
RECON_MUSCL, RECON_PC = 0, 1
RIEMANN_HLL, RIEMANN_RUSANOV, RIEMANN_ROE = 0, 1, 2

DT_FIXED = 0      # one dt for every step
DT_ADAPTIVE = 1   # dt from the CFL number at each step's start state (recorded)
DT_GIVEN = 2      # a caller-given dt per step

G_REF = 9.80665  # reference include/constants.hpp:4

_dp = C.POINTER(C.c_double)


class GridDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("ng", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("g", C.c_double),
        ("bc_kind", C.c_int32),
    ]


class SweError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


# every symbol include/swe_hip.h declares, with (restype, argtypes)
_SIGS = {
    "swe_hip_abi_version": (C.c_int, []),
    "swe_hip_set_scheme": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "swe_hip_get_scheme": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_int)]),
    "swe_hip_error_string": (C.c_char_p, [C.c_int]),
    "swe_hip_last_error": (C.c_char_p, [C.c_void_p]),
    "swe_hip_last_status": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    "swe_hip_set_step_fusion": (C.c_int, [C.c_void_p, C.c_int]),
    "swe_hip_get_step_fusion": (C.c_int, [C.c_void_p]),
    "swe_hip_device_count": (C.c_int, []),
    "swe_hip_create": (C.c_int, [C.POINTER(GridDesc), C.c_int, C.POINTER(C.c_void_p)]),
    "swe_hip_destroy": (None, [C.c_void_p]),
    "swe_hip_upload": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp]),
    "swe_hip_download": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "swe_hip_upload_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p]),
    "swe_hip_download_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "swe_hip_stable_dt": (C.c_int, [C.c_void_p, C.c_double, _dp, _dp]),
    "swe_hip_step": (C.c_int, [C.c_void_p, C.c_double, C.c_int64]),
    "swe_hip_step_adaptive": (C.c_int, [C.c_void_p, C.c_double, C.c_int64, _dp]),
    "swe_hip_step_adaptive_noalloc": (C.c_int, [C.c_void_p, C.c_double, C.c_int64,
                                                C.c_void_p]),
    "swe_hip_rhs": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "swe_hip_set_tile_rows": (C.c_int, [C.c_void_p, C.c_int]),
    "swe_hip_get_tile_rows": (C.c_int, [C.c_void_p]),
    "swe_hip_set_stage_tile_rows": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "swe_hip_get_stage_tile_rows": (C.c_int, [C.c_void_p, C.c_int]),
    "swe_hip_set_flat_bottom": (C.c_int, [C.c_void_p, C.c_int]),
    "swe_hip_get_flat_bottom": (C.c_int, [C.c_void_p]),
    "swe_hip_set_precision": (C.c_int, [C.c_void_p, C.c_int]),
    "swe_hip_get_precision": (C.c_int, [C.c_void_p]),
    "swe_hip_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "swe_hip_set_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "swe_hip_time_steps": (C.c_int, [C.c_void_p, C.c_double, C.c_int64, _dp, _dp]),
    "swe_hip_sync": (C.c_int, [C.c_void_p]),
    "swe_hip_batch_create": (C.c_int, [C.POINTER(GridDesc), C.c_int32, C.c_int,
                                       C.POINTER(C.c_void_p)]),
    "swe_hip_batch_destroy": (None, [C.c_void_p]),
    "swe_hip_batch_last_error": (C.c_char_p, [C.c_void_p]),
    "swe_hip_batch_count": (C.c_int, [C.c_void_p]),
    "swe_hip_batch_upload": (C.c_int, [C.c_void_p, C.c_int32, _dp, _dp, _dp, _dp]),
    "swe_hip_batch_download": (C.c_int, [C.c_void_p, C.c_int32, _dp, _dp, _dp]),
    "swe_hip_batch_set_dt": (C.c_int, [C.c_void_p, _dp]),
    "swe_hip_batch_stable_dt": (C.c_int, [C.c_void_p, C.c_double, _dp]),
    "swe_hip_batch_step": (C.c_int, [C.c_void_p, C.c_int64]),
    "swe_hip_batch_step_adaptive": (C.c_int, [C.c_void_p, C.c_double, C.c_int64, _dp]),
    "swe_hip_batch_nonfinite": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "swe_hip_batch_set_launch": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "swe_hip_batch_get_launch": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "swe_hip_batch_autotune": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "swe_hip_batch_set_precision": (C.c_int, [C.c_void_p, C.c_int]),
    "swe_hip_batch_set_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "swe_hip_batch_sync": (C.c_int, [C.c_void_p]),
    "swe_hip_adj_create": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.POINTER(C.c_void_p)]),
    "swe_hip_adj_create_budget": (C.c_int, [C.c_void_p, C.c_int64, C.c_double,
                                            C.POINTER(C.c_void_p)]),
    "swe_hip_adj_plan": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                   C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)]),
    "swe_hip_adj_plan_budget": (C.c_int, [C.c_int64, C.c_int64, C.c_double,
                                          C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                          C.POINTER(C.c_int64)]),
    "swe_hip_adj_destroy": (None, [C.c_void_p]),
    "swe_hip_adj_last_error": (C.c_char_p, [C.c_void_p]),
    "swe_hip_adj_forward": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, _dp]),
    "swe_hip_adj_backward": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "swe_hip_adj_backward_device": (C.c_int, [C.c_void_p] + [C.c_void_p] * 7),
    "swe_hip_adj_set_launch": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "swe_hip_adj_schedule": (C.c_int, [C.c_int64, C.c_int32, C.POINTER(C.c_int64),
                                       C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    "swe_hip_adj_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                   _dp, _dp]),
}

_lib = None


def lib():
    """Load libswe_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        # One HIP runtime per process.  torch ships its own libamdhip64; if this library
        # (linked against the system ROCm) initializes HIP first, torch's runtime then finds
        # no GPU.  Importing torch first makes the library resolve to torch's runtime, so
        # both share the device (and torch tensors can be passed as device pointers).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m swe_hip.build` "
                               "(or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def device_count() -> int:
    return lib().swe_hip_device_count()


def _p(a: np.ndarray):
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise TypeError("arrays must be C-contiguous float64")
    return a.ctypes.data_as(_dp)


class HipSolver:
    """One grid on one GPU.  Mirrors the operations of the reference's backend classes
    (FastHLLMUSCLBathyOpenMPSolver / FastHLLMUSCLBathyCUDASolver) over raw arrays."""

    def __init__(self, nx, ny, dx, dy, g=G_REF, bc=BC_REFLECTING, ng=2, device=0):
        self.nx, self.ny, self.ng = int(nx), int(ny), int(ng)
        self.shape = (self.nx + 2 * self.ng, self.ny + 2 * self.ng)
        self.desc = GridDesc(self.nx, self.ny, self.ng, float(dx), float(dy), float(g), int(bc))
        self._ctx = C.c_void_p()
        L = lib()
        self._check(L.swe_hip_create(C.byref(self.desc), int(device), C.byref(self._ctx)))

    @classmethod
    def from_case(cls, case, device=0):
        return cls(case["nx"], case["ny"], case["dx"], case["dy"], case["g"], case["bc"],
                   case["ng"], device)

    def _check(self, rc):
        if rc != SWE_OK:
            L = lib()
            detail = ""
            if self._ctx:
                detail = (L.swe_hip_last_error(self._ctx) or b"").decode()
            raise SweError(rc, L.swe_hip_error_string(rc).decode() + (f": {detail}" if detail else ""))

    def close(self):
        if self._ctx:
            lib().swe_hip_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _arr(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"expected shape {self.shape}, got {a.shape}")
        return a

    def upload(self, h, hu, hv, B):
        h, hu, hv, B = (self._arr(a) for a in (h, hu, hv, B))
        self._check(lib().swe_hip_upload(self._ctx, _p(h), _p(hu), _p(hv), _p(B)))

    def download(self):
        out = [np.empty(self.shape) for _ in range(3)]
        self._check(lib().swe_hip_download(self._ctx, *(_p(o) for o in out)))
        return out

    def upload_device(self, h, hu, hv, B):
        """torch float64 tensors on this context's device, shape (nx+2ng, ny+2ng)."""
        for t in (h, hu, hv, B):
            if tuple(t.shape) != self.shape or not t.is_contiguous():
                raise ValueError("device tensors must be contiguous with the grid shape")
        self._check(lib().swe_hip_upload_device(self._ctx, h.data_ptr(), hu.data_ptr(),
                                                hv.data_ptr(), B.data_ptr()))

    def download_device(self, h, hu, hv):
        for t in (h, hu, hv):
            if tuple(t.shape) != self.shape or not t.is_contiguous():
                raise ValueError("device tensors must be contiguous with the grid shape")
        self._check(lib().swe_hip_download_device(self._ctx, h.data_ptr(), hu.data_ptr(),
                                                  hv.data_ptr()))

    def stable_dt(self, cfl):
        dt, ms = C.c_double(), C.c_double()
        self._check(lib().swe_hip_stable_dt(self._ctx, float(cfl), C.byref(dt), C.byref(ms)))
        return dt.value, ms.value

    def step(self, dt, nsteps=1):
        self._check(lib().swe_hip_step(self._ctx, float(dt), int(nsteps)))

    def step_adaptive(self, cfl, nsteps=1):
        hist = np.zeros(max(int(nsteps), 1))
        # kept for callers that catch SweError: the dts up to a non-finite stop
        self.last_dt_hist = hist[:nsteps]
        self._check(lib().swe_hip_step_adaptive(self._ctx, float(cfl), int(nsteps), _p(hist)))
        return hist[:nsteps]

    @property
    def step_fusion(self):
        """Rows per tile of the fused one-launch step (0: three launches per step)."""
        return lib().swe_hip_get_step_fusion(self._ctx)

    @step_fusion.setter
    def step_fusion(self, rows):
        self._check(lib().swe_hip_set_step_fusion(self._ctx, int(rows)))

    def last_status(self):
        """(steps started, failing stage or 0) of the last stepping call: after
        SWE_ERR_NONFINITE the call stopped at that stage of that step (1-based)."""
        step, stage = C.c_int64(), C.c_int()
        self._check(lib().swe_hip_last_status(self._ctx, C.byref(step), C.byref(stage)))
        return step.value, stage.value

    def step_adaptive_noalloc(self, cfl, nsteps=1, dt_hist=None):
        """swe_hip_step_adaptive_noalloc: no device allocation; dt_hist is None or a float64
        device tensor of at least nsteps entries (written on the device)."""
        ptr = None
        if dt_hist is not None:
            if str(dt_hist.dtype) != "torch.float64" or not dt_hist.is_cuda or \
                    dt_hist.numel() < int(nsteps) or not dt_hist.is_contiguous():
                raise TypeError("dt_hist must be a contiguous float64 device tensor of nsteps")
            ptr = C.c_void_p(dt_hist.data_ptr())
        self._check(lib().swe_hip_step_adaptive_noalloc(self._ctx, float(cfl), int(nsteps), ptr))

    def rhs(self):
        out = [np.empty(self.shape) for _ in range(3)]
        self._check(lib().swe_hip_rhs(self._ctx, *(_p(o) for o in out)))
        return out

    @property
    def tile_rows(self):
        return lib().swe_hip_get_tile_rows(self._ctx)

    @tile_rows.setter
    def tile_rows(self, v):
        self._check(lib().swe_hip_set_tile_rows(self._ctx, int(v)))

    @property
    def flat_bottom(self):
        """True when the stage launches skip B (a bottom of +0.0 everywhere, mode on)."""
        return lib().swe_hip_get_flat_bottom(self._ctx) == 1

    def set_flat_bottom(self, enabled):
        """Mode of the flat-bottom kernels: True (default) uses them when B is +0.0
        everywhere, False always reads B."""
        self._check(lib().swe_hip_set_flat_bottom(self._ctx, 1 if enabled else 0))

    @property
    def stage_tile_rows(self):
        """Tile heights of the three stage launches (stage 1, 2, 3)."""
        return tuple(lib().swe_hip_get_stage_tile_rows(self._ctx, s) for s in (1, 2, 3))

    @stage_tile_rows.setter
    def stage_tile_rows(self, rows):
        for s, v in zip((1, 2, 3), rows):
            self._check(lib().swe_hip_set_stage_tile_rows(self._ctx, s, int(v)))

    @property
    def precision(self):
        return lib().swe_hip_get_precision(self._ctx)

    @precision.setter
    def precision(self, mode):
        self._check(lib().swe_hip_set_precision(self._ctx, int(mode)))

    def set_scheme(self, scheme: int, recon: int = RECON_MUSCL, riemann: int = RIEMANN_HLL):
        """Scheme family of step/rhs (SCHEME_*): the optimized scheme, or the reference's
        modular serial solver with the given reconstruction and Riemann solver."""
        self._check(lib().swe_hip_set_scheme(self._ctx, int(scheme), int(recon), int(riemann)))

    def get_scheme(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        self._check(lib().swe_hip_get_scheme(self._ctx, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def set_graph(self, enabled: bool):
        self._check(lib().swe_hip_set_graph(self._ctx, int(bool(enabled))))

    def set_stream(self, stream_handle: int | None):
        self._check(lib().swe_hip_set_stream(self._ctx, stream_handle or None))

    def time_steps(self, dt, nsteps):
        st = (C.c_double * 3)()
        tot = C.c_double()
        self._check(lib().swe_hip_time_steps(self._ctx, float(dt), int(nsteps), st, C.byref(tot)))
        return list(st), tot.value

    def sync(self):
        self._check(lib().swe_hip_sync(self._ctx))


class HipBatch:
    """An ensemble of independent grids advanced together (swe_hip_batch_* in the C ABI).

    cases: sequence of workload dicts (nx, ny, dx, dy, g, bc) -- all with ng = 2 and the
    same g and bc."""

    def __init__(self, cases, device=0):
        self.count = len(cases)
        descs = (GridDesc * self.count)()
        self.shapes = []
        for i, c in enumerate(cases):
            descs[i] = GridDesc(int(c["nx"]), int(c["ny"]), int(c.get("ng", 2)), float(c["dx"]),
                                float(c["dy"]), float(c["g"]), int(c["bc"]))
            self.shapes.append((int(c["nx"]) + 4, int(c["ny"]) + 4))
        self._b = C.c_void_p()
        self._check(lib().swe_hip_batch_create(descs, self.count, int(device), C.byref(self._b)))

    def _check(self, rc):
        if rc != SWE_OK:
            L = lib()
            detail = (L.swe_hip_batch_last_error(self._b) or b"").decode() if self._b else ""
            raise SweError(rc, L.swe_hip_error_string(rc).decode() + (f": {detail}" if detail else ""))

    def close(self):
        if self._b:
            lib().swe_hip_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, i, h, hu, hv, B):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (h, hu, hv, B)]
        for a in arrs:
            if a.shape != self.shapes[i]:
                raise ValueError(f"instance {i}: expected {self.shapes[i]}, got {a.shape}")
        self._check(lib().swe_hip_batch_upload(self._b, int(i), *(_p(a) for a in arrs)))

    def download(self, i):
        out = [np.empty(self.shapes[i]) for _ in range(3)]
        self._check(lib().swe_hip_batch_download(self._b, int(i), *(_p(o) for o in out)))
        return out

    def set_dt(self, dt):
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        assert dt.shape == (self.count,)
        self._check(lib().swe_hip_batch_set_dt(self._b, _p(dt)))

    def stable_dt(self, cfl):
        out = np.zeros(self.count)
        self._check(lib().swe_hip_batch_stable_dt(self._b, float(cfl), _p(out)))
        return out

    def step(self, nsteps=1):
        self._check(lib().swe_hip_batch_step(self._b, int(nsteps)))

    def step_adaptive(self, cfl, nsteps=1):
        hist = np.zeros((max(int(nsteps), 1), self.count))
        # kept for callers that catch SweError: the dts up to each instance's stop
        self.last_dt_hist = hist[:nsteps]
        self._check(lib().swe_hip_batch_step_adaptive(self._b, float(cfl), int(nsteps), _p(hist)))
        return hist[:nsteps]

    def nonfinite(self):
        f = (C.c_int32 * self.count)()
        self._check(lib().swe_hip_batch_nonfinite(self._b, f))
        return np.array(f[:], dtype=np.int32)

    def set_launch(self, tile_rows, waves_per_cu=0):
        self._check(lib().swe_hip_batch_set_launch(self._b, int(tile_rows), int(waves_per_cu)))

    def get_launch(self):
        t, w = C.c_int(), C.c_int()
        self._check(lib().swe_hip_batch_get_launch(self._b, C.byref(t), C.byref(w)))
        return t.value, w.value

    def autotune(self):
        t, w = C.c_int(), C.c_int()
        self._check(lib().swe_hip_batch_autotune(self._b, C.byref(t), C.byref(w)))
        return t.value, w.value

    def set_precision(self, mode):
        self._check(lib().swe_hip_batch_set_precision(self._b, int(mode)))

    def set_graph(self, enabled: bool):
        self._check(lib().swe_hip_batch_set_graph(self._b, int(bool(enabled))))

    def sync(self):
        self._check(lib().swe_hip_batch_sync(self._b))


def adjoint_schedule(nsteps: int, nsnap: int):
    """(forward steps advanced in all, most advances of one step, schedule actions) of the
    binomial checkpoint schedule; host-only, needs no GPU."""
    adv, reps, acts = C.c_int64(), C.c_int32(), C.c_int64()
    rc = lib().swe_hip_adj_schedule(int(nsteps), int(nsnap), C.byref(adv), C.byref(reps),
                                    C.byref(acts))
    if rc != SWE_OK:
        raise SweError(rc, lib().swe_hip_error_string(rc).decode())
    return adv.value, reps.value, acts.value


def adjoint_plan_budget(nsteps: int, n_total: int, max_bytes: float):
    """(recorded steps K, checkpoint interval m, stored states) of the storage plan
    HipAdjoint(..., budget_bytes=max_bytes) picks; K = 0 means binomial checkpoints.
    Host-only, needs no GPU."""
    k, m, st = C.c_int64(), C.c_int64(), C.c_int64()
    rc = lib().swe_hip_adj_plan_budget(int(nsteps), int(n_total), float(max_bytes), C.byref(k),
                                       C.byref(m), C.byref(st))
    if rc != SWE_OK:
        raise SweError(rc, lib().swe_hip_error_string(rc).decode())
    return k.value, m.value, st.value


class HipAdjoint:
    """Discrete adjoint of `nsteps` SSP-RK3 steps of a HipSolver (swe_hip_adj_* in the C ABI).

    forward() runs the steps from the solver's current state, storing binomial checkpoints
    (nsnap stored states, U^0 included; -1 = every step with its stage states; 0 = the
    budget plan below for 60% of the free device memory) or, with budget_bytes, the plan
    with the fewest recomputed steps in that much memory (the last steps' stage records
    plus periodic checkpoints)
    and leaves U^N in the solver; backward(lam) takes dL/dU^N and returns
    dL/dU^0 and dL/dB with dt frozen from the forward pass.  One backward per forward."""

    def __init__(self, solver: HipSolver, nsteps: int, nsnap: int = 0,
                 budget_bytes: float | None = None):
        self.solver = solver
        self.nsteps = int(nsteps)
        self._a = C.c_void_p()
        if budget_bytes is not None:
            self._check(lib().swe_hip_adj_create_budget(solver._ctx, self.nsteps,
                                                        float(budget_bytes), C.byref(self._a)))
        else:
            self._check(lib().swe_hip_adj_create(solver._ctx, self.nsteps, int(nsnap),
                                                 C.byref(self._a)))

    def _check(self, rc):
        if rc != SWE_OK:
            L = lib()
            detail = (L.swe_hip_adj_last_error(self._a) or b"").decode() if self._a else ""
            raise SweError(rc, L.swe_hip_error_string(rc).decode() + (f": {detail}" if detail else ""))

    def close(self):
        if self._a:
            lib().swe_hip_adj_destroy(self._a)
            self._a = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def forward(self, dt=None, cfl=None, dts=None):
        """Exactly one of: dt (fixed), cfl (adaptive, recorded), dts (given per step).
        Returns the dt sequence used."""
        hist = np.zeros(self.nsteps)
        if dts is not None:
            hist[:] = np.asarray(dts, dtype=np.float64)
            self._check(lib().swe_hip_adj_forward(self._a, DT_GIVEN, 0.0, _p(hist)))
        elif cfl is not None:
            self._check(lib().swe_hip_adj_forward(self._a, DT_ADAPTIVE, float(cfl), _p(hist)))
        else:
            self._check(lib().swe_hip_adj_forward(self._a, DT_FIXED, float(dt), _p(hist)))
        return hist

    def backward(self, lam_h, lam_hu, lam_hv):
        lam = [self.solver._arr(a) for a in (lam_h, lam_hu, lam_hv)]
        out = [np.empty(self.solver.shape) for _ in range(4)]
        self._check(lib().swe_hip_adj_backward(self._a, *(_p(a) for a in lam),
                                               *(_p(o) for o in out)))
        return out  # dL/dh0, dL/dhu0, dL/dhv0, dL/dB

    def backward_device(self, lam_h, lam_hu, lam_hv, g_h, g_hu, g_hv, g_b=None):
        """torch float64 device tensors (contiguous, grid shape)."""
        ts = [lam_h, lam_hu, lam_hv, g_h, g_hu, g_hv] + ([g_b] if g_b is not None else [])
        for t in ts:
            if tuple(t.shape) != self.solver.shape or not t.is_contiguous():
                raise ValueError("device tensors must be contiguous with the grid shape")
        self._check(lib().swe_hip_adj_backward_device(
            self._a, lam_h.data_ptr(), lam_hu.data_ptr(), lam_hv.data_ptr(), g_h.data_ptr(),
            g_hu.data_ptr(), g_hv.data_ptr(), g_b.data_ptr() if g_b is not None else None))

    def set_launch(self, block_waves: int = 4, rows: int = 0, y_rows: int = 0):
        """Launch shape of the adjoint stages: block_waves 2, 4 or 8 with `rows` rows per wave
        (fused kernel), or 0 for the two-kernel path (`rows` x faces per wave, `y_rows`
        rows per wave of the y kernel).  0 keeps a value."""
        self._check(lib().swe_hip_adj_set_launch(self._a, int(block_waves), int(rows),
                                                 int(y_rows)))

    def plan(self):
        """Storage plan: recorded steps, checkpoint interval, stored states, their bytes and
        the forward work the backward recomputes (in steps)."""
        k, m, st, by, rc = C.c_int64(), C.c_int64(), C.c_int64(), C.c_double(), C.c_double()
        self._check(lib().swe_hip_adj_plan(self._a, C.byref(k), C.byref(m), C.byref(st),
                                           C.byref(by), C.byref(rc)))
        return {"record_steps": k.value, "ckpt_every": m.value, "stored_states": st.value,
                "store_bytes": by.value, "recomputed_steps": rc.value}

    def info(self):
        n, r, f, b = C.c_int32(), C.c_int32(), C.c_double(), C.c_double()
        self._check(lib().swe_hip_adj_info(self._a, C.byref(n), C.byref(r), C.byref(f),
                                           C.byref(b)))
        return {"nsnap": n.value, "records": bool(r.value), "forward_ms": f.value,
                "backward_ms": b.value}
