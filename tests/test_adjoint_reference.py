"""Pins of the FP64 autograd reference used by the adjoint tests
(tests/adjoint_reference.py), on the CPU.

1. Its forward step equals the oracle's propose_phi followed by
   implicit_diffusion (pearl_serial.cpp:961-1094) on the 48x40 fixture:
   phases and carbon to rounding, mu to 1e-8 J/mol (the oracle computes
   mu' = mu_ref + (c' - ic) / chi from the PCG iterate, the reference the
   exact solve; the cementite chi of about 5e-9 amplifies the solve's
   rounding).
2. The stencil-PCG solve with implicit differentiation (solve="cg") and
   checkpointed segments give the dense-solve gradient.
3. Autograd of the reference agrees with central differences of the same
   torch model (24 steps; h = 1e-6 relative for two pair energies and a mobility,
   1e-3 for the diffusivities, whose influence on the loss over 24 steps
   is about 1e-3 of sigma's), so the reference gradient is the derivative
   of the model it restates, projection included.
"""
import numpy as np
import pytest
import torch

import adjoint_reference as R
from conftest import load_golden


def _case(steps):
    g = load_golden("small_48x40")
    pd = dict(g["pd"])
    pd["cg_rtol"], pd["cg_atol"] = 1e-15, 1e-30
    mu_ref = float(g["sc0"][9])
    dts = np.full(steps, float(g["dt_probe"]))
    H0 = float((1.0 - g["phi0"][2]).sum() * pd["dy"] / pd["nx"])
    return g, pd, mu_ref, dts, 1e6, H0 - 1e-10


@pytest.mark.parametrize("solve", ["dense", "cg"])
def test_reference_step_equals_oracle_step(O, solve):
    g, pd, mu_ref, dts, _, _ = _case(1)
    dt = float(dts[0])
    P = O.params_from_dict(pd)
    _, phi_t, st = O.propose_phi(P, g["phi0"], g["mu0"], dt)
    assert st == 0
    d = O.implicit_diffusion(P, phi_t, g["c0"], g["mu0"], dt, mu_ref)
    assert d["ok"]
    th = [torch.tensor(v, dtype=torch.float64) for v in R.theta_of(pd)]
    pt, cn, mn = R.step(torch.tensor(g["phi0"]), torch.tensor(g["c0"]), torch.tensor(g["mu0"]), th, pd, mu_ref, dt,
                        solve)
    assert np.max(np.abs(pt.numpy() - phi_t)) <= 1e-15
    assert np.max(np.abs(cn.numpy() - d["carbon"])) <= 1e-15 * np.max(np.abs(d["carbon"]))
    assert np.max(np.abs(mn.numpy() - d["mu"])) <= 1e-8


def test_cg_solve_gradient_equals_dense():
    g, pd, mu_ref, dts, w, ht = _case(6)
    args = (pd, mu_ref, g["phi0"], g["c0"], g["mu0"], dts, g["phi0"], w, ht)
    Ld, gd, *_ = R.loss_and_grad(*args)
    Lc, gc, *_ = R.loss_and_grad(*args, solve="cg", every=4)
    assert abs(Lc - Ld) <= 1e-14 * abs(Ld)
    assert np.all(np.abs(gc - gd) <= 1e-12 * np.abs(gd))


def test_reference_gradient_matches_its_finite_differences():
    g, pd, mu_ref, dts, w, ht = _case(24)

    def loss(q):
        return R.loss_and_grad(q, mu_ref, g["phi0"], g["c0"], g["mu0"], dts, g["phi0"], w, ht)[0]

    _, gr, *_ = R.loss_and_grad(pd, mu_ref, g["phi0"], g["c0"], g["mu0"], dts, g["phi0"], w, ht)
    for name, rh in (("sigma_ag", 1e-6), ("sigma_tg", 1e-6), ("tau_tg", 1e-6), ("D0", 1e-3), ("D1", 1e-3)):
        k = R.THETA.index(name)
        base = R.theta_of(pd)[k]
        h = rh * base
        a, b = dict(pd), dict(pd)
        a[name] = base + h
        b[name] = base - h
        fd = (loss(a) - loss(b)) / (2 * h)
        assert abs(fd - gr[k]) <= 1e-6 * abs(gr[k]), (name, fd, gr[k])
