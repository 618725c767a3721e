"""A domain far above the production size: 16384x8192 cells (134 M cells,
about 22.5 GB of device state), 128 lamella pairs, initialized on the
device (pearl_fill_phases, pearl_smooth, pearl_init_composition).  The
oracle cannot run at this size, so the checks are properties that hold at
any size:

1. propose_phi is per-cell work with an exact max reduction, and x is
   periodic: shifting the state by s columns shifts the trial phases by s
   columns and leaves the maximum change unchanged, bit for bit.
2. Five accepted steps keep the state valid (valid_state :1096-1116), the
   phase fractions on the simplex, and the carbon ledger closed: the
   flux-form update conserves carbon to rounding, far inside the reference's
   mass test (1e-11 relative, :1152)."""
import importlib
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
NX, NY, PAIRS, STRIP = 16384, 8192, 128, 1024


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_large_domain_invariance_and_conservation(pkg, tmp_path):
    inputs = importlib.import_module("pearl_amd.inputs")
    config = importlib.import_module("pearl_amd.config")
    path = inputs.colony_input(str(tmp_path / "large.in"), NX, NY, PAIRS, STRIP, 1e-6, spacing_noise=0.05, seed=3,
                               overrides={"smooth_steps": 10, "smooth_min_steps": 10})
    cfg = config.read_config(path)
    ops, n_ops = pkg.fill_ops(cfg)
    say("large domain", NX, NY, "cells", NX * NY)
    with pkg.Solver(cfg.solver_params(), cfg.initial_mu) as s:
        s.fill_phases(ops, n_ops)
        rep = s.smooth(cfg.smooth_steps, cfg.smooth_min_steps, cfg.smooth_energy_rtol, cfg.smooth_max_phase_change)
        mu0, carbon0, valid = s.init_composition(pkg.COMPOSITION_MU, cfg.initial_mu)
        say("  initialized on the device: smoothing steps", rep["steps"], "valid", valid)
        assert valid and rep["steps"] == cfg.smooth_steps
        phi, c, mu, _ = s.get_state()
        sc = [0.0, cfg.dt, 0.0, carbon0, 0.0, 0.0, 0, 0, 0]
        s.set_state(phi, c, mu, sc)
        dt = cfg.dt
        ch1, phit1 = s.propose_phi(dt)
        shift = 1237
        s.set_state(np.roll(phi, shift, axis=-1), np.roll(c, shift, axis=-1), np.roll(mu, shift, axis=-1), sc)
        ch2, phit2 = s.propose_phi(dt)
        say("  propose: max change", ch1, ch2)
        assert ch1 == ch2 and ch1 > 0
        assert np.array_equal(np.roll(phit1, shift, axis=-1), phit2)
        del phit1, phit2
        # five accepted steps from the unshifted state
        s.set_state(phi, c, mu, sc)
        del phi, c, mu
        for _ in range(5):
            s.advance(1.0)
        d = s.diagnostics()
        st = s.scalars()
        say("  after 5 steps: t", st.time, "rejected", st.rejected, "phi-sum error", d["phi_sum_error"])
        assert st.step == 5 and s.valid_state()
        assert d["phi_sum_error"] <= 1e-14
        assert abs(s.total_carbon() - carbon0) <= 1e-13 * abs(carbon0)
