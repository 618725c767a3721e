"""The oracle against the reference compiled from source (oracle/_ref),
live: more states than the committed fixtures, the smoothing pass, and the
blocked reduction order over the whole smoke trajectory.  Skipped where the
reference build is absent."""
import os
import tempfile

import numpy as np
import pytest

from conftest import load_golden, load_pkg

import oracle as O

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")


def inputs_mod():
    load_pkg()
    import importlib
    return importlib.import_module("pearl_amd.inputs")


@pytest.mark.parametrize("uc,pairs,nx,ny", [(5.0, 2, 40, 24), (15.0, 3, 60, 30), (40.0, 2, 32, 40)])
def test_trajectories_bitwise(uc, pairs, nx, ny):
    I = inputs_mod()
    p = I.colony_input(tempfile.mktemp(suffix=".in"), nx, ny, pairs, ny // 3, 6e-10, undercooling=uc,
                       overrides={"smooth_steps": 15, "smooth_min_steps": 15})
    R = O.Ref(p)
    R.initialize()
    phi, c, mu, sc = R.get_state()
    P = O.params_from_dict(R.pd)
    S = O.OracleState(phi, c, mu, sc[:9])
    for _ in range(25):
        info, err = R.advance(6e-10)
        rc, oi = O.advance(P, S, sc[9], 6e-10)
        if err:
            assert rc != 0
            break
        assert rc == 0
        assert list(info[:7]) == [oi.accepted_dt, oi.max_phase_change, oi.linear_residual, oi.closure,
                                  oi.energy_change, oi.cg_iterations, oi.retries]
    phi1, c1, mu1, sc1 = R.get_state()
    assert np.array_equal(phi1, S.phi) and np.array_equal(c1, S.carbon) and np.array_equal(mu1, S.mu)


@pytest.mark.parametrize("cg_max", [2000, 9, 8])
def test_restarts_and_iteration_limit_bitwise(cg_max):
    """Solves whose recursive residual passes while the true residual does
    not (cg_rtol 1e-16, cg_atol at the input's smallest value), so they
    restart from the true residual; with cg_max 9 and 8 some end unconverged
    and the step is retried.  The oracle's advance equals the reference's,
    step info and fields."""
    I = inputs_mod()
    p = I.colony_input(tempfile.mktemp(suffix=".in"), 40, 24, 2, 8, 6e-10,
                       overrides={"smooth_steps": 10, "smooth_min_steps": 10, "cg_rtol": "1e-16",
                                  "cg_atol": "1e-300", "cg_max": str(cg_max)})
    R = O.Ref(p)
    R.initialize()
    phi, c, mu, sc = R.get_state()
    P = O.params_from_dict(R.pd)
    S = O.OracleState(phi, c, mu, sc[:9])
    retries = 0
    for _ in range(12):
        info, err = R.advance(6e-10)
        rc, oi = O.advance(P, S, sc[9], 6e-10)
        assert not err and rc == 0
        assert list(info[:7]) == [oi.accepted_dt, oi.max_phase_change, oi.linear_residual, oi.closure,
                                  oi.energy_change, oi.cg_iterations, oi.retries]
        retries += oi.retries
    if cg_max < 10:
        assert retries > 0
    phi1, c1, mu1, _ = R.get_state()
    assert np.array_equal(phi1, S.phi) and np.array_equal(c1, S.carbon) and np.array_equal(mu1, S.mu)


def test_smoothing_bitwise():
    I = inputs_mod()
    p = I.colony_input(tempfile.mktemp(suffix=".in"), 48, 32, 2, 10, 1e-9)
    R = O.Ref(p)
    R.initialize()  # includes the 100-step smoothing of the sharp labels
    phi_r, _, _, _ = R.get_state()
    # restate: sharp labels -> oracle smoothing
    load_pkg()
    import importlib
    initial = importlib.import_module("pearl_amd.initial")
    config = importlib.import_module("pearl_amd.config")
    cfg = config.read_config(p)
    phi = initial.sharp_labels(cfg.nx, cfg.ny, cfg.filling)
    P = O.params_from_dict(R.pd)
    rc, rep = O.smooth(P, phi, cfg.smooth_steps, cfg.smooth_min_steps, cfg.smooth_energy_rtol,
                       cfg.smooth_max_phase_change)
    assert rc == 0
    assert np.array_equal(phi, phi_r)


def test_forced_rejections_match_reference():
    """Host-forceable failure paths give the reference's messages."""
    g = load_golden("small_48x40")
    I = inputs_mod()
    p = I.colony_input(tempfile.mktemp(suffix=".in"), 48, 40, 2, 10, 1.5e-9,
                       overrides={"smooth_steps": 20, "smooth_min_steps": 20})
    for ctl, msg in [({"max_phase_change": 1e-12}, "step rejected below dt_min: phase increment"),
                     ({"cg_max": 0}, None)]:
        R = O.Ref(p)
        R.set_controls(**ctl)
        R.set_state(g["phi0"], g["c0"], g["mu0"], g["sc0"])
        info, err = R.advance(1.5e-9)
        pd = dict(g["pd"])
        pd.update(ctl)
        P = O.params_from_dict(pd)
        S = O.OracleState(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
        rc, oi = O.advance(P, S, float(g["sc0"][9]), 1.5e-9)
        assert err is not None and rc in (2, 3)
        prefix = "step rejected below dt_min: " if rc == 2 else "step retry limit: "
        assert err == prefix + O.WHY[oi.why]
        if msg:
            assert err == msg


def test_blocked_order_full_smoke_trajectory_within_bar():
    """Over the whole 256x128 smoke trajectory (0.015 us), the oracle in the
    HIP path's blocked summation order takes the same accepted steps,
    rejections and PCG iterations as the reference, with final fields inside
    the FP64 bar (|dmu| <= 1e-5 J/mol, |dphi| <= 1e-10)."""
    g = load_golden("colony_256x128_full")
    P = O.params_from_dict(g["pd"], red_mode=O.RED_BLOCKED, red_k=1)
    S = O.OracleState(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
    rc, st = O.run(P, S, float(g["sc0"][9]), float(g["end_time"]), float(g["output_dt"]), 10 ** 9)
    assert rc == 0
    assert st["taken"] == int(g["taken"]) and st["cg_total"] == int(g["cg_total"])
    assert S.st.rejected == int(g["sc1"][8])
    assert np.abs(S.phi - g["phi1"]).max() <= 1e-10
    assert np.abs(S.mu - g["mu1"]).max() <= 1e-5
