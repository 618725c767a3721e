"""Host-side parameter handling on the CPU (no device): the pair interface
energies in the C-ABI parameter block, the adjoint's gradient order, and the
torch harness's parameter vector."""
import importlib

import numpy as np

from conftest import load_golden


def test_pair_sigma_fills_the_parameter_block(pkg):
    g = load_golden("small_48x40")
    p = pkg.params_from_dict(g["pd"])
    assert list(p.pair_sigma) == [0.0, 0.0, 0.0]  # unset: the reference's single sigma
    p = pkg.params_from_dict(dict(g["pd"], sigma_at=0.5, sigma_ag=0.4, sigma_tg=0.6))
    assert list(p.pair_sigma) == [0.5, 0.4, 0.6]
    # the C layout: the four ints end at byte 256, where the 8-aligned array starts
    import ctypes
    assert pkg.Params.pair_sigma.offset == pkg.Params.phase_only.offset + 4 == 256
    assert ctypes.sizeof(pkg.Params) == 280


def test_adjoint_gradient_order(pkg):
    assert pkg.ADJOINT_GRAD[:7] == ("sigma", "tau_at", "tau_ag", "tau_tg", "D0", "D1", "D2")
    assert pkg.ADJOINT_GRAD[7:] == pkg.PAIR_SIGMA_KEYS


def test_fixed_step_problem_parameters(pkg):
    TA = importlib.import_module("pearl_amd.torch_adjoint")
    g = load_golden("small_48x40")
    pd = dict(g["pd"])
    prob = TA.FixedStepProblem(pd, float(g["sc0"][9]), (g["phi0"], g["c0"], g["mu0"], None), [1e-12] * 3,
                               g["phi0"], w_front=2.0, every=2)
    th = prob.theta_values()
    assert th[:3] == [pd["sigma"]] * 3 and th[3:6] == [pd["tau_at"], pd["tau_ag"], pd["tau_tg"]]
    q = prob.params_at([1.0, 2.0, 3.0] + th[3:])
    assert (q["sigma_at"], q["sigma_ag"], q["sigma_tg"]) == (1.0, 2.0, 3.0) and q["sigma"] == pd["sigma"]
    # the harness reorders pearl_adjoint's gradient into its parameter order
    grad10 = np.arange(10.0)
    assert list(grad10[TA._FROM_ADJOINT]) == [7, 8, 9, 1, 2, 3, 4, 5, 6]
    # loss of given phases: 1/2 |phi - target|^2 + 1/2 w (H - H_target)^2
    # (phi + 1e-3 - phi is 1e-3 only to the rounding of phi, ~1e-13 relative)
    phi = g["phi0"].copy()
    phi[2] += 1e-3
    H = TA.front_height(phi, pd["dy"])
    want = 0.5 * 3.0 * 0.0 + 0.5 * (1e-3) ** 2 * phi[2].size + 0.5 * 2.0 * (H - prob.front_target) ** 2
    assert abs(prob.loss_of_phases(phi) - want) <= 1e-12 * want
