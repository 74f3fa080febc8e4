"""SWE_PRECISION_FAST: reciprocal-based velocities.  Not bitwise; the bound here is what the
mode is documented to deliver on smooth flows (few-ulp differences per step), and the test
also records that it is NOT within the 1e-12 parity bar on the noisy dam break."""
import numpy as np
import pytest

import oracle
from swe_hip import workloads as W

pytestmark = pytest.mark.gpu


def rel_linf(a, b, ng=2):
    I = (slice(ng, -ng), slice(ng, -ng))
    return np.abs(a[I] - b[I]).max() / max(np.abs(b[I]).max(), 1e-300)


def run(hip, c, precision, steps, dt):
    with hip.HipSolver.from_case(c) as s:
        s.upload(c["h"], c["hu"], c["hv"], c["B"])
        s.precision = precision
        s.step(dt, steps)
        return s.download()


def test_fast_mode_close_on_smooth_flow(hip):
    c = W.gauss_pulse(256, g=9.81)
    gr = oracle.grid(256, 256, 2, c["dx"], c["dy"], g=9.81)
    dt, _ = oracle.stable_dt(gr, c["h"], c["hu"], c["hv"], 0.4)
    exact = run(hip, c, hip.PRECISION_EXACT, 300, dt)
    fast = run(hip, c, hip.PRECISION_FAST, 300, dt)
    assert rel_linf(fast[0], exact[0]) < 1e-12
    assert rel_linf(fast[1], exact[1]) < 1e-10
    assert rel_linf(fast[2], exact[2]) < 1e-10


def test_fast_mode_error_on_noisy_dam_break(hip):
    c = W.dam_break(1024, steps=200)
    gr = oracle.grid(1024, 1024, 2, c["dx"], c["dy"])
    dt, _ = oracle.stable_dt(gr, c["h"], c["hu"], c["hv"], 0.45)
    exact = run(hip, c, hip.PRECISION_EXACT, 200, dt)
    fast = run(hip, c, hip.PRECISION_FAST, 200, dt)
    errs = [rel_linf(f, e) for f, e in zip(fast, exact)]
    # small but not bitwise: the mode is opt-in and never used for the parity claims
    assert errs[0] < 1e-10 and 0.0 < max(errs) < 1e-5, errs
