"""The step loop between output times on the device (pearl_run with the
output tolerance of pearl_set_stop_tolerance) against the reference's loop
shape (pearl_serial.cpp:2296-2313): one advance per step, and an output
once time >= next_output - eps, eps = 1e-12*max(end_time, output_dt).

Bar: bitwise equal states, clocks, step counts and rejection counts at every
output time, and equal per-output step counts.  Output intervals of a few
steps and of a fraction of a step are both covered, so runs end both on
dt clipped to the output time and on the step limit."""
import sys

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(pkg, g):
    s = pkg.Solver(g["pd"], float(g["sc0"][9]))
    s.set_state(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
    return s


def reference_loop(s, end_time, output_dt, outputs):
    """advance() per step, outputs as in the reference's main loop."""
    eps = 1e-12 * max(end_time, output_dt)
    nxt = output_dt
    marks = []
    while len(marks) < outputs and s.scalars().time < end_time:
        s.advance(min(end_time, nxt))
        t = s.scalars().time
        if t >= nxt - eps or t >= end_time - eps:
            sc = s.scalars()
            marks.append((sc.time, sc.step, sc.rejected, s.get_state()[:3]))
            nxt += output_dt
    return marks


def device_loop(s, end_time, output_dt, outputs, chunk):
    eps = 1e-12 * max(end_time, output_dt)
    s.set_stop_tolerance(eps)
    nxt = output_dt
    marks = []
    while len(marks) < outputs and s.scalars().time < end_time:
        s.run(min(end_time, nxt), chunk)
        t = s.scalars().time
        if t >= nxt - eps or t >= end_time - eps:
            sc = s.scalars()
            marks.append((sc.time, sc.step, sc.rejected, s.get_state()[:3]))
            nxt += output_dt
    return marks


@pytest.mark.parametrize("name,output_steps,chunk", [("small_48x40", 7.3, 512), ("colony_256x128", 3.0, 2),
                                                      ("small_48x40", 0.4, 512)])
def test_run_between_outputs_equals_advance_loop(pkg, name, output_steps, chunk):
    g = load_golden(name)
    dt = float(g["pd"]["dt_max"])
    output_dt = output_steps * dt
    end_time = 1.0
    say("run loop", name, output_steps, chunk)
    with make(pkg, g) as a:
        ref = reference_loop(a, end_time, output_dt, 6)
    with make(pkg, g) as b:
        got = device_loop(b, end_time, output_dt, 6, chunk)
    assert len(ref) == len(got) == 6
    for (t0, s0, r0, f0), (t1, s1, r1, f1) in zip(ref, got):
        say(f"  t {t0:.6e} step {s0} rejected {r0}")
        assert t0 == t1 and s0 == s1 and r0 == r1
        for x, y in zip(f0, f1):
            assert np.array_equal(x, y)


def test_stop_tolerance_rejects_invalid(pkg):
    g = load_golden("small_48x40")
    with make(pkg, g) as s:
        with pytest.raises(pkg.PearlError):
            s.set_stop_tolerance(-1.0)
        with pytest.raises(pkg.PearlError):
            s.set_stop_tolerance(float("nan"))
