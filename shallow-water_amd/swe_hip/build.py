"""Build libswe_hip.so (hipcc, gfx950) in-tree.

    python -m swe_hip.build            # from shallow-water_amd/

The library is written next to this file so it travels with the source tree.
Flags:
  -ffp-contract=off   no a*b+c -> fma contraction: keeps the reference's rounding order
                      (the CPU solvers are built without FMA), so results are bitwise equal.
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed-FP32 instructions (none are needed by this FP64 code).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
CSRC = os.path.join(PKG_ROOT, "csrc")
REPO = os.path.dirname(PKG_ROOT)
LIB = os.path.join(HERE, "libswe_hip.so")

SOURCES = ["swe_stage.hip", "swe_api.hip", "swe_batch_api.hip", "swe_adjoint.hip",
           "swe_adjoint_api.hip", "swe_modular.hip"]
ARCH = os.environ.get("SWE_HIP_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)


def build(verbose: bool = False, extra=None) -> str:
    """Compile each source to an object in parallel (the stage kernels' file dominates), then
    link the shared library."""
    import tempfile

    inc = ["-I", os.path.join(REPO, "include")]
    compile_flags = [f for f in HIPCC_FLAGS if f not in ("-shared",)]
    with tempfile.TemporaryDirectory(prefix="swe_hip_build_") as tmpdir:
        objs, procs = [], []
        for src in SOURCES:
            obj = os.path.join(tmpdir, os.path.splitext(src)[0] + ".o")
            objs.append(obj)
            procs.append((src, _run([hipcc(), *compile_flags, *(extra or []), *inc, "-c",
                                     "-o", obj, os.path.join(CSRC, src)], verbose)))
        failed = []
        for src, pr in procs:
            out, _ = pr.communicate()
            if pr.returncode != 0:
                failed.append((src, pr.returncode, out))
        if failed:
            for src, rc, out in failed:
                sys.stderr.write(f"--- {src} ({rc})\n{out}")
            raise RuntimeError(f"hipcc failed on {', '.join(f[0] for f in failed)}")
        tmp = LIB + ".tmp"
        link = _run([hipcc(), *HIPCC_FLAGS, *(extra or []), "-o", tmp, *objs], verbose)
        out, _ = link.communicate()
        if link.returncode != 0:
            sys.stderr.write(out)
            raise RuntimeError(f"hipcc link failed ({link.returncode})")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
