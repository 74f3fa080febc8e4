"""The cheaper instruction forms of the stage kernel are bitwise equal to the forms they
replace (CPU, IEEE-754 double arithmetic; see tests/exact_forms.c):

* the split division's last step with the remainder negated instead of a copysign of the
  numerator (swe_face.h div_numerator), for reciprocal estimates within a few ulps;
* the median form of 0.5 * minmod (swe_face.h half_minmod_split) against
  omp_kernels.hpp:45-48 on differences of values inside the kernel's range guard.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_exact_forms(tmp_path):
    exe = tmp_path / "exact_forms"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe),
                    os.path.join(HERE, "exact_forms.c"), "-lm"], check=True)
    out = subprocess.run([str(exe), "2000000"], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.split() == ["div", "0", "median", "0"]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_exact_forms_checker_catches_a_wrong_form(tmp_path):
    """The harness is not vacuous: a median with min and max swapped in the clamp (wrong
    for two negative slopes) or the reference division without its copysign (wrong sign
    of a -0 quotient) is detected."""
    src = open(os.path.join(HERE, "exact_forms.c")).read()
    broken = src.replace("fmax(fmin(a, b), fmin(fmax(a, b), 0.0))",
                         "fmax(fmin(a, b), fmax(fmin(a, b), 0.0))")
    broken = broken.replace("copysign(fma(fma(-d, q, num), r, q), num)",
                            "fma(fma(-d, q, num), r, q)")
    assert broken != src
    c = tmp_path / "broken.c"
    c.write_text(broken)
    exe = tmp_path / "broken"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(c), "-lm"],
                   check=True)
    out = subprocess.run([str(exe), "200000"], capture_output=True, text=True)
    assert out.returncode == 1
    _, div_bad, _, med_bad = out.stdout.split()
    assert int(div_bad) > 0 and int(med_bad) > 0
