"""The C-ABI library: it loads, exports every function include/swe_hip.h declares, and
validates descriptors before touching a device.  No compute calls (runs without a GPU)."""
import ctypes as C
import os
import re

import pytest

import swe_hip

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "swe_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(swe_hip_\w+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_abi():
    names = declared_functions()
    assert "swe_hip_create" in names and "swe_hip_step" in names and len(names) >= 15


def test_library_exports_every_declared_symbol():
    L = swe_hip.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_every_symbol():
    assert set(declared_functions()) == set(swe_hip._SIGS)


def test_abi_version_and_error_strings():
    L = swe_hip.lib()
    assert L.swe_hip_abi_version() == 1
    for code in range(0, 7):
        assert L.swe_hip_error_string(code)
    assert L.swe_hip_error_string(99) == b"unknown error"
    assert L.swe_hip_last_error(None) == b""


@pytest.mark.parametrize("desc", [
    (1, 10, 2, 0.1, 0.1, 9.8, 0),      # nx < ng
    (10, 10, 1, 0.1, 0.1, 9.8, 0),     # ng < 2 (MUSCL needs 2 ghost layers)
    (10, 10, 2, 0.0, 0.1, 9.8, 0),     # dx <= 0
    (10, 10, 2, 0.1, 0.1, -1.0, 0),    # g <= 0
    (10, 10, 2, 0.1, 0.1, 9.8, 7),     # unknown boundary kind
    (50000, 50000, 2, 0.1, 0.1, 9.8, 0),  # beyond 32-bit cell indexing
])
def test_create_rejects_bad_descriptors(desc):
    L = swe_hip.lib()
    ctx = C.c_void_p()
    rc = L.swe_hip_create(C.byref(swe_hip.GridDesc(*desc)), 0, C.byref(ctx))
    assert rc == swe_hip.SWE_ERR_INVALID
    assert not ctx


def test_null_arguments_are_rejected():
    L = swe_hip.lib()
    assert L.swe_hip_create(None, 0, None) == swe_hip.SWE_ERR_INVALID
    assert L.swe_hip_step(None, 0.1, 1) == swe_hip.SWE_ERR_INVALID
    assert L.swe_hip_download(None, None, None, None) == swe_hip.SWE_ERR_INVALID
    L.swe_hip_destroy(None)  # no-op


def test_no_device_is_reported_without_gpu():
    if swe_hip.device_count() > 0:
        pytest.skip("a GPU is visible")
    L = swe_hip.lib()
    ctx = C.c_void_p()
    rc = L.swe_hip_create(C.byref(swe_hip.GridDesc(16, 16, 2, 0.1, 0.1, 9.8, 0)), 0,
                          C.byref(ctx))
    assert rc == swe_hip.SWE_ERR_NO_DEVICE


def test_solver_raises_without_gpu():
    if swe_hip.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(swe_hip.SweError) as e:
        swe_hip.HipSolver(16, 16, 0.1, 0.1)
    assert e.value.code == swe_hip.SWE_ERR_NO_DEVICE
