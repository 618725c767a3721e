"""The hand-written adjoint (pearl_adjoint) and its torch.autograd.Function
harness against FP64 autograd references and central finite differences.

Small case: the 48x40 coupled fixture, 24 fixed steps of the fixture's
probe dt, the initial phases as target, the front term on.  The reference
is tests/adjoint_reference.py with the dense solve (the step restated in
torch, differentiated by autograd).  PCG tolerance tightened to 1e-15
relative so the forward and the exact-solve adjoint agree to rounding.

Colony case: the parameter-recovery problem at 512x256 (4 lamella pairs,
15 K; torch_adjoint.colony_inverse_problem) over its first 200 steps,
checkpointed every 50.  The reference is the same restatement with the
stencil PCG solve, differentiated implicitly, run in checkpointed segments
on the GPU in FP64.

Parameters: the interface energy of each phase pair (sigma_at, sigma_ag,
sigma_tg), the mobility coefficients and the diffusivities.  The small case
runs twice: with the reference's single sigma (pair energies equal; the
forward is the reference's step) and with three different pair energies.
The single-sigma gradient (pearl_adjoint's grad[0]) is the sum of the pair
gradients and is checked against a central difference of `sigma` itself.

Bars: 1e-7 relative against the dense reference, 1e-8 against the colony
reference; 1e-6 relative against central differences of the GPU forward.
The difference step is 1e-6 relative for sigma and the mobilities.  The
loss depends weakly on the diffusivities over 24 steps (theta*dL/dtheta / L
is 1e-3 to 1e-4 there, against about 1 for the others), so a 1e-6 step
moves the loss by 1e-9 to 1e-10 of itself, where the forward's rounding
noise dominates the quotient; their step is 1e-3 relative (truncation about
1e-8, measured on the reference), and the 1e-6 quotients are printed beside
it."""
import sys

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu
STEPS = 24
PAIRS = {"sigma_at": 0.55, "sigma_ag": 0.42, "sigma_tg": 0.61}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def setup(pairs=None):
    g = load_golden("small_48x40")
    pd = dict(g["pd"])
    pd.update(pairs or {})
    pd["cg_rtol"], pd["cg_atol"] = 1e-15, 1e-30
    mu_ref = float(g["sc0"][9])
    dts = np.full(STEPS, float(g["dt_probe"]))
    H0 = float((1.0 - g["phi0"][2]).sum() * pd["dy"] / pd["nx"])
    return g, pd, mu_ref, dts, g["phi0"].copy(), 1e6, H0 - 1e-10


def gpu_loss(pkg, g, pd, mu_ref, dts, target, w, h_target):
    with pkg.Solver(pd, mu_ref) as s:
        s.set_state(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
        for dt in dts:
            s.step_fixed(dt)
        phi, _, _, _ = s.get_state()
    H = float((1.0 - phi[2]).sum() * pd["dy"] / pd["nx"])
    return 0.5 * float(((phi - target) ** 2).sum()) + 0.5 * w * (H - h_target) ** 2


def gpu_adjoint(pkg, g, pd, mu_ref, dts, target, w, h_target, every):
    with pkg.Solver(pd, mu_ref) as s:
        s.set_state(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
        return s.adjoint(dts, target, every=every, w_front=w, front_target=h_target)


def by_name(pkg, grad):
    return dict(zip(pkg.ADJOINT_GRAD, grad))


@pytest.fixture(scope="module")
def adj(pkg):
    g, pd, mu_ref, dts, target, w, ht = setup()
    say("adjoint, every=4")
    return gpu_adjoint(pkg, g, pd, mu_ref, dts, target, w, ht, 4)


@pytest.mark.parametrize("pairs", [None, PAIRS], ids=["single-sigma", "pair-sigma"])
def test_adjoint_matches_autograd_reference(pkg, pairs):
    import adjoint_reference as R
    g, pd, mu_ref, dts, target, w, ht = setup(pairs)
    say("adjoint vs autograd", "pairs" if pairs else "single sigma")
    L, grad = gpu_adjoint(pkg, g, pd, mu_ref, dts, target, w, ht, 4)
    gd = by_name(pkg, grad)
    Lr, gr, *_ = R.loss_and_grad(pd, mu_ref, g["phi0"], g["c0"], g["mu0"], dts, target, w, ht)
    say("loss", L, Lr)
    for k, name in enumerate(R.THETA):
        say(f"  {name:8s} adjoint {gd[name]: .12e} autograd {gr[k]: .12e} rel {abs(gd[name] - gr[k]) / abs(gr[k]):.2e}")
    assert abs(L - Lr) <= 1e-9 * abs(Lr)
    for k, name in enumerate(R.THETA):
        assert abs(gd[name] - gr[k]) <= 1e-7 * abs(gr[k])
    assert abs(gd["sigma"] - (gd["sigma_at"] + gd["sigma_ag"] + gd["sigma_tg"])) <= 1e-14 * abs(gd["sigma"])


H_REL = {"sigma": 1e-6, "sigma_at": 1e-6, "sigma_ag": 1e-6, "sigma_tg": 1e-6, "tau_at": 1e-6, "tau_ag": 1e-6,
         "tau_tg": 1e-6, "D0": 1e-3, "D1": 1e-3, "D2": 1e-3}


def central_difference(pkg, g, pd, mu_ref, dts, target, w, ht, name, rel_h):
    base = float(pd.get(name) or pd["sigma"])  # a pair energy left unset is sigma
    h = rel_h * base
    vals = []
    for sgn in (1, -1):
        q = dict(pd)
        q[name] = base + sgn * h
        vals.append(gpu_loss(pkg, g, q, mu_ref, dts, target, w, ht))
    return (vals[0] - vals[1]) / (2 * h)


def test_adjoint_matches_finite_differences(pkg, adj):
    g, pd, mu_ref, dts, target, w, ht = setup()
    L, grad = adj
    gd = by_name(pkg, grad)
    for name in pkg.ADJOINT_GRAD:
        k = pkg.ADJOINT_GRAD.index(name)
        fd = central_difference(pkg, g, pd, mu_ref, dts, target, w, ht, name, H_REL[name])
        rel = abs(grad[k] - fd) / abs(fd)
        extra = ""
        if H_REL[name] != 1e-6:
            fd6 = central_difference(pkg, g, pd, mu_ref, dts, target, w, ht, name, 1e-6)
            extra = f"   (h=1e-6: fd {fd6: .12e} rel {abs(grad[k] - fd6) / abs(fd6):.2e})"
        say(f"  {name:8s} adjoint {grad[k]: .12e} fd(h={H_REL[name]:.0e}) {fd: .12e} rel {rel:.2e}{extra}")
        assert rel <= 1e-6


def test_adjoint_checkpoint_interval_invariant(pkg, adj):
    g, pd, mu_ref, dts, target, w, ht = setup()
    L4, g4 = adj
    for every in (1, 7, STEPS):
        L, gr = gpu_adjoint(pkg, g, pd, mu_ref, dts, target, w, ht, every)
        assert L == L4
        assert np.allclose(gr, g4, rtol=1e-12, atol=0)


# ---- colony case and the autograd.Function harness -------------------------

@pytest.fixture(scope="module")
def colony(pkg):
    import importlib
    TA = importlib.import_module("pearl_amd.torch_adjoint")
    say("colony inverse problem 512x256, 200 steps")
    prob, factors, _ = TA.colony_inverse_problem(steps=200, every=50)
    return TA, prob, factors


def test_colony_adjoint_matches_reference(pkg, colony):
    import torch
    import adjoint_reference as R
    TA, prob, _ = colony
    th = prob.theta0().requires_grad_(True)
    L = TA.phase_field_loss(th, prob)
    L.backward()
    g = th.grad.numpy()
    phi, c, mu, _ = prob.state
    say("torch reference (stencil PCG, checkpointed)")
    Lr, gr, phr, *_ = R.loss_and_grad(prob.params, prob.mu_reference, phi, c, mu, prob.dts, prob.phi_target,
                                      prob.w_front, prob.front_target, solve="cg", device="cuda", every=50)
    torch.cuda.empty_cache()
    for k, name in enumerate(TA.THETA):
        say(f"  {name:7s} adjoint {g[k]: .12e} reference {gr[k]: .12e} rel {abs(g[k] - gr[k]) / abs(gr[k]):.2e}")
    assert abs(L.item() - Lr) <= 1e-9 * abs(Lr)
    assert float(np.max(np.abs(prob.final_phases(prob.theta0().tolist()) - phr))) <= 1e-11
    for k in range(len(TA.THETA)):
        assert abs(g[k] - gr[k]) <= 1e-8 * abs(gr[k])


def test_autograd_function_harness(pkg, colony):
    """PhaseFieldLoss: the loss equals the forward-only loss, backward gives
    the adjoint gradient (scaled by the incoming gradient), and a small
    descent step along it lowers the loss by what the gradient predicts."""
    import torch
    TA, prob, _ = colony
    th = prob.theta0().requires_grad_(True)
    L = TA.phase_field_loss(th, prob)
    (3.0 * L).backward()
    L0, g0 = prob.loss_and_grad(prob.theta0().tolist())
    assert L.item() == L0
    assert abs(L.item() - prob.loss(prob.theta0().tolist())) <= 1e-12 * L0
    assert np.array_equal(th.grad.numpy(), 3.0 * g0)
    # descent in log-parameters: d theta_k = -a theta_k^2 g_k; the first-order
    # prediction of the change is -a sum (theta_k g_k)^2
    t = prob.theta0().numpy()
    a = 1e-4 * L0 / float(np.sum((t * g0) ** 2))
    t1 = t - a * t * t * g0
    dL = prob.loss(t1.tolist()) - L0
    pred = -a * float(np.sum((t * g0) ** 2))
    say(f"  descent: dL {dL:.6e} predicted {pred:.6e}")
    assert dL < 0 and abs(dL - pred) <= 1e-2 * abs(pred)


def test_run_fixed_equals_step_fixed(pkg):
    """pearl_run_fixed (all steps under the device controller, one call) is
    bitwise the same as one pearl_step_fixed call per size."""
    g, pd, mu_ref, dts, *_ = setup(PAIRS)
    dts = dts * np.linspace(0.5, 1.0, dts.size)
    with pkg.Solver(pd, mu_ref) as a, pkg.Solver(pd, mu_ref) as b:
        for s in (a, b):
            s.set_state(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
        for dt in dts:
            a.step_fixed(float(dt))
        b.run_fixed(dts)
        for x, y in zip(a.get_state()[:3], b.get_state()[:3]):
            assert np.array_equal(x, y)
        assert a.scalars().time == b.scalars().time and a.scalars().step == b.scalars().step


def test_fixed_entry_points_reject_bad_arguments_and_leave_fixed_mode(pkg):
    g, pd, mu_ref, dts, target, *_ = setup()
    with pkg.Solver(pd, mu_ref) as s:
        s.set_state(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
        for bad in ([], [0.0], [1e-12, float("nan")], [-1e-12]):
            with pytest.raises(pkg.PearlError):
                s.run_fixed(bad)
        with pytest.raises(pkg.PearlError):
            s.adjoint(dts[:2], target, every=0)
        with pytest.raises(pkg.PearlError):
            s.adjoint(dts[:2], target, every=1, w_front=-1.0)
        # after fixed steps the context advances under the adaptive
        # controller exactly like a fresh context from the same state
        s.run_fixed(dts[:3])
        phi, c, mu, sc = s.get_state()
        for _ in range(4):
            s.advance(1e9)
        a = s.get_state()
    with pkg.Solver(pd, mu_ref) as f:
        f.set_state(phi, c, mu, sc)
        for _ in range(4):
            f.advance(1e9)
        b = f.get_state()
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)
    assert a[3].time == b[3].time and a[3].step == b[3].step
    with pkg.Solver(dict(pd, phase_only=1), mu_ref) as s:
        s.set_state(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
        with pytest.raises(pkg.PearlError):
            s.run_fixed(dts[:2])
