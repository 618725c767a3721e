"""The oracle restatement against the golden fixtures written by the
reference itself (tests/golden/make_golden.py).  Sequential-order results
must be bitwise equal to the reference."""
import numpy as np
import pytest

from conftest import load_golden

CASES = ["small_48x40", "phase_only_48x40", "window_32x40", "window_alpha_32x40", "colony_256x128", "tiny_3x2",
         "ragged_257x9"]

SHIFT_ERRORS = {"unsafe shift: removed strip contains gamma": 6, "invalid shift": 8}


@pytest.mark.parametrize("name", CASES)
def test_single_functions_bitwise(O, name):
    g = load_golden(name)
    P = O.params_from_dict(g["pd"])
    mu_ref = float(g["sc0"][9])
    ch, phit, st = O.propose_phi(P, g["phi0"], g["mu0"], float(g["dt_probe"]))
    assert st == 0
    assert ch == g["propose_change"]
    assert np.array_equal(phit, g["propose_phi"])
    d = O.implicit_diffusion(P, g["propose_phi"], g["c0"], g["mu0"], float(g["dt_probe"]), mu_ref)
    assert d["status"] == 0
    assert d["ok"] == g["diff_ok"] and d["it"] == g["diff_it"]
    assert d["residual"] == g["diff_residual"] and d["closure"] == g["diff_closure"]
    if d["ok"]:
        assert np.array_equal(d["carbon"], g["diff_carbon"])
        assert np.array_equal(d["mu"], g["diff_mu"])
    e, st = O.reduced_energy(P, g["phi0"], g["c0"], mu_ref)
    assert st == 0 and e == g["energy0"]
    assert O.total_carbon(P, g["c0"]) == g["carbon0"]
    assert O.valid_state(P, g["phi0"], g["c0"], g["mu0"]) == g["valid0"]


@pytest.mark.parametrize("name", CASES)
def test_trajectory_bitwise(O, name):
    g = load_golden(name)
    P = O.params_from_dict(g["pd"])
    mu_ref = float(g["sc0"][9])
    S = O.OracleState(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
    for k in range(int(g["steps"])):
        rc, info = O.advance(P, S, mu_ref, float(g["end_time"]))
        assert rc == 0
        ref = g["step_infos"][k]
        got = [info.accepted_dt, info.max_phase_change, info.linear_residual, info.closure,
               info.energy_change, info.cg_iterations, info.retries]
        assert got == list(ref[:7]), (k, got, ref)
    assert np.array_equal(S.phi, g["phi1"])
    assert np.array_equal(S.carbon, g["c1"])
    assert np.array_equal(S.mu, g["mu1"])
    assert S.scalars() == list(g["sc1"][:9])


@pytest.mark.parametrize("name", ["window_32x40", "window_alpha_32x40"])
def test_moving_window_bitwise(O, name):
    g = load_golden(name)
    P = O.params_from_dict(g["pd"])
    S = O.OracleState(g["phi1"], g["c1"], g["mu1"], g["sc1"][:9])
    rc, sh, front = O.maybe_shift(P, S, 1, *[float(v) for v in g["window"]])
    err = str(g["shift_error"])
    if err:
        assert rc == SHIFT_ERRORS[err]
        return
    assert rc == 0 and sh == g["shifted"] == 1
    assert np.array_equal(S.phi, g["phi2"])
    assert np.array_equal(S.carbon, g["c2"])
    assert np.array_equal(S.mu, g["mu2"])
    assert S.scalars() == list(g["sc2"][:9])
