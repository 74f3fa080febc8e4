"""CPU builds with AddressSanitizer and UndefinedBehaviorSanitizer (g++/gcc only; no GPU
code is instrumented):

* the oracle (oracle/swe_oracle.c, swe_modular.c) and the dual-number tangent
  (oracle/swe_tangent.cpp), driven by tests/sanitize_driver.cpp over odd grid shapes, both
  boundary kinds, ghost widths 2 and 3, wet and wet/dry states and every modular
  reconstruction x Riemann combination;
* the host application (shallow-water_amd/host) in its GPU-free modes: configuration
  parsing and validation (good and bad files) and the initial-state NetCDF output.

A sanitizer finding aborts the program or prints "runtime error" on stderr; both fail the
test."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
       "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


def have_sanitizers():
    if not shutil.which("gcc") or not shutil.which("g++"):
        return False
    r = subprocess.run(["g++", *SAN, "-x", "c++", "-", "-o", os.devnull],
                       input="int main(){return 0;}", capture_output=True, text=True)
    return r.returncode == 0


pytestmark = pytest.mark.skipif(not have_sanitizers(), reason="no sanitizer runtime")


def clean(r):
    return r.returncode == 0 and "runtime error" not in r.stderr and "Sanitizer" not in r.stderr


def test_oracle_and_tangent_under_asan_ubsan(tmp_path):
    o = str(tmp_path)
    flags = SAN + ["-ffp-contract=off", "-fopenmp", "-Wall"]
    for src, obj, cc, std in (("oracle/swe_oracle.c", "o1.o", "gcc", "-std=c11"),
                              ("oracle/swe_modular.c", "o2.o", "gcc", "-std=c11"),
                              ("oracle/swe_tangent.cpp", "t.o", "g++", "-std=c++17"),
                              ("tests/sanitize_driver.cpp", "d.o", "g++", "-std=c++17")):
        subprocess.run([cc, std, *flags, "-c", os.path.join(REPO, src), "-o",
                        os.path.join(o, obj)], check=True)
    exe = os.path.join(o, "driver")
    subprocess.run(["g++", *flags, *(os.path.join(o, x) for x in ("o1.o", "o2.o", "t.o", "d.o")),
                    "-o", exe, "-lm"], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, env=dict(ENV, OMP_NUM_THREADS="2"),
                       timeout=600)
    assert clean(r), r.stdout[-2000:] + r.stderr[-4000:]
    assert "0 failures" in r.stdout


def test_host_app_under_asan_ubsan(tmp_path):
    lib = os.path.join(REPO, "shallow-water_amd", "swe_hip", "libswe_hip.so")
    if not os.path.exists(lib):
        pytest.skip("libswe_hip.so not built")
    build = str(tmp_path / "build")
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "shallow-water_amd", "host"),
                    f"BUILD={build}",
                    "CXXFLAGS=-std=c++17 -Wall -Wextra -Wno-unused-parameter " + " ".join(SAN),
                    "LDFLAGS=" + " ".join(SAN)], check=True)
    app = os.path.join(build, "swe_hip_app")
    env = dict(ENV, LD_LIBRARY_PATH=os.path.dirname(lib), ASAN_OPTIONS="detect_leaks=0")
    import test_host_app as H

    cases = [("--print-config", {}), ("--print-config", {"mesh__Nx": "40.5"}),
             ("--print-config", {"backend__type": '"Vulkan"'}),
             ("--print-config", {"time__cfl": 1.5})]
    for args, over in cases:
        r = subprocess.run([app, args, H.config(tmp_path, **over)], capture_output=True,
                           text=True, env=env, timeout=120)
        assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr[-4000:]
        assert r.returncode in (0, 1)
    for name in ("GaussInitial__GaussHill", "DamBreakRadial__Flat", "c1_gauss__Flat"):
        out = str(tmp_path / f"{name}.nc")
        cfg = H.config(tmp_path, **H.setup_overrides(H.SETUPS[name]))
        r = subprocess.run([app, "--initial-state", out, cfg], capture_output=True, text=True,
                           env=env, timeout=120)
        assert clean(r), r.stderr[-4000:]
    bad = tmp_path / "bad.toml"
    bad.write_text("[mesh]\nLx = [1, 2]\n")
    r = subprocess.run([app, "--print-config", str(bad)], capture_output=True, text=True,
                       env=env, timeout=120)
    assert r.returncode == 1 and "runtime error" not in r.stderr and "Sanitizer" not in r.stderr
