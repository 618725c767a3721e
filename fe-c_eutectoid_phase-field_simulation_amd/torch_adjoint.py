"""torch.autograd.Function over the hand-written adjoint (pearl_adjoint).

The differentiated map is a fixed sequence of always-taken steps
(propose_phi, pearl_serial.cpp:961-1002, then implicit_diffusion,
:1010-1094) from a given state; the dt sequence is frozen (the adaptive
controller's accept/reject and clipping are not differentiated).  The loss
is the L2 mismatch of the final phase fields plus a weighted front-height
term:

    L = 1/2 sum (phi - phi_target)^2 + 1/2 w_front (H - H_target)^2,
    H = dy / nx * sum over cells of (1 - phi_gamma)   (mean colony height)

Parameters theta = (sigma_at, sigma_ag, sigma_tg, tau_at, tau_ag, tau_tg,
D_alpha, D_theta, D_gamma): the interface energy of each phase pair
(ferrite-cementite, ferrite-austenite, cementite-austenite; the reference has
one sigma for all three, and equal pair energies reproduce it), the three
pair mobility coefficients (the reference's tau) and the three carbon
diffusivities.

`FixedStepProblem` holds everything but theta.  `PhaseFieldLoss.apply(theta,
problem)` runs pearl_adjoint once: the forward with checkpoints every
`problem.every` steps, the loss, and the reverse sweep with recomputation
per segment.  It returns the loss as a 0-d tensor and keeps the gradient
for backward, which scales it by the incoming gradient.  Everything runs on
the GPU through the C ABI; theta itself may live on any device.

In a process that uses torch on the GPU, import torch before the first call
into the library: libpearl_gpu.so then binds to the HIP runtime that torch
loaded (this module imports torch first).  Loaded the other way round, the
process holds two HIP runtimes and torch finds no device.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import ADJOINT_GRAD, Solver

THETA = ("sigma_at", "sigma_ag", "sigma_tg", "tau_at", "tau_ag", "tau_tg", "D0", "D1", "D2")
_FROM_ADJOINT = [ADJOINT_GRAD.index(k) for k in THETA]


def front_height(phi: np.ndarray, dy: float) -> float:
    """Mean colony height: dy / nx * sum (1 - phi_gamma), summed exactly
    (math.fsum), like pearl_adjoint's compensated sum."""
    nx = phi.shape[-1]
    return math.fsum((1.0 - phi[2]).ravel().tolist()) * dy / nx


class FixedStepProblem:
    """The loss of a frozen-dt run, as a function of theta.

    params: solver parameters (dict with the C-ABI keys); state: (phi[3,ny,nx],
    carbon, mu, scalars or None) at the start; dts: the step sizes;
    phi_target[3,ny,nx]; front_target: target mean colony height (default:
    the height of phi_target); w_front >= 0; every: checkpoint interval of
    the adjoint."""

    def __init__(self, params, mu_reference, state, dts, phi_target, front_target=None, w_front=0.0, every=100,
                 device=0):
        self.params = dict(params)
        self.mu_reference = float(mu_reference)
        self.state = state
        self.dts = np.ascontiguousarray(dts, dtype=np.float64)
        self.phi_target = np.ascontiguousarray(phi_target, dtype=np.float64)
        self.front_target = (front_height(self.phi_target, self.params["dy"]) if front_target is None
                             else float(front_target))
        self.w_front = float(w_front)
        self.every = int(every)
        self.device = int(device)
        if self.dts.ndim != 1 or self.dts.size < 1:
            raise ValueError("dts must be a non-empty sequence of step sizes")
        if self.every < 1:
            raise ValueError("every must be >= 1")

    def params_at(self, theta) -> dict:
        th = [float(v) for v in theta]
        if len(th) != len(THETA):
            raise ValueError(f"theta has {len(th)} entries, expected {len(THETA)}")
        p = dict(self.params)
        p.update(zip(THETA, th))
        return p

    def theta_values(self) -> list:
        """THETA at the problem's parameters (a pair energy left 0 is sigma)."""
        return [float(self.params.get(k) or self.params["sigma"]) if k.startswith("sigma_") else float(self.params[k])
                for k in THETA]

    def theta0(self, dtype=torch.float64, device="cpu") -> torch.Tensor:
        return torch.tensor(self.theta_values(), dtype=dtype, device=device)

    def _solver(self, theta):
        s = Solver(self.params_at(theta), self.mu_reference, device=self.device)
        phi, c, mu, sc = self.state
        s.set_state(phi, c, mu, sc)
        return s

    def final_phases(self, theta) -> np.ndarray:
        """Forward only: the phases after the fixed steps."""
        with self._solver(theta) as s:
            s.run_fixed(self.dts)
            return s.get_state()[0]

    def loss_of_phases(self, phi: np.ndarray) -> float:
        h = front_height(phi, self.params["dy"]) - self.front_target
        return 0.5 * math.fsum(((phi - self.phi_target) ** 2).ravel().tolist()) + 0.5 * self.w_front * h * h

    def loss(self, theta) -> float:
        return self.loss_of_phases(self.final_phases(theta))

    def loss_and_grad(self, theta):
        """(loss, dL/dtheta in THETA order) from one pearl_adjoint call."""
        with self._solver(theta) as s:
            loss, g = s.adjoint(self.dts, self.phi_target, every=self.every, w_front=self.w_front,
                                front_target=self.front_target)
        return loss, g[_FROM_ADJOINT]


class PhaseFieldLoss(torch.autograd.Function):
    """loss = PhaseFieldLoss.apply(theta, problem); loss.backward() fills
    theta.grad with the adjoint gradient."""

    @staticmethod
    def forward(ctx, theta, problem):
        if theta.dim() != 1 or theta.numel() != len(THETA):
            raise ValueError(f"theta must be a vector of {len(THETA)} entries")
        loss, grad = problem.loss_and_grad(theta.detach().cpu().double().tolist())
        ctx.save_for_backward(torch.as_tensor(grad, dtype=theta.dtype, device=theta.device))
        return theta.new_tensor(loss)

    @staticmethod
    def backward(ctx, grad_out):
        (grad,) = ctx.saved_tensors
        return grad_out * grad, None


def phase_field_loss(theta: torch.Tensor, problem: FixedStepProblem) -> torch.Tensor:
    return PhaseFieldLoss.apply(theta, problem)


def colony_inverse_problem(nx=512, ny=256, pairs=4, undercooling=15.0, steps=2000, every=100, seed=20261203,
                           factor_seed=20261203, device=0, directory=None):
    """The parameter-recovery case: a lamellar colony (pairs ferrite/cementite
    pairs, colony height ny/8, +/-5 % spacing) at the given undercooling.

    1. An adaptive run (advance) of `steps` accepted steps with the
       unscaled parameters fixes the dt sequence.
    2. The target: the same fixed steps with the ferrite/austenite and
       cementite/austenite interface energies (sigma_ag, sigma_tg) and
       mobility coefficients (tau_ag, tau_tg) scaled by seeded factors in
       [0.8, 1.2].
    3. The loss starts from the unscaled parameters; w_front = nx / dy^2, so
       that a one-cell shift of the front weighs about like the phase
       mismatch of that shift.
    The PCG tolerances are tightened to 1e-15 relative (1e-30 absolute): the
    adjoint differentiates the exact solve, and the forward then agrees with
    it to rounding.  Returns (problem, factors, config)."""
    import random
    import tempfile

    from . import inputs
    tmp = directory or tempfile.mkdtemp(prefix="pearl_inverse_")
    _, cfg, phi, c, mu = inputs.colony_workload(nx, ny, pairs, ny // 8, seed, undercooling, tmp, moving_window=0)
    params = cfg.solver_params()
    params["cg_rtol"], params["cg_atol"] = 1e-15, 1e-30
    with Solver(params, cfg.initial_mu, device=device) as s:
        s.set_state(phi, c, mu, [0.0, cfg.dt, 0.0, 0.0, 0.0, 0.0, 0, 0, 0])
        m0 = s.total_carbon()
        scalars = [0.0, cfg.dt, 0.0, m0, 0.0, 0.0, 0, 0, 0]
        s.set_state(phi, c, mu, scalars)
        dts = np.array([s.advance(1.0).accepted_dt for _ in range(steps)])
    rng = random.Random(factor_seed)
    factors = {k: rng.uniform(0.8, 1.2) for k in ("sigma_ag", "sigma_tg", "tau_ag", "tau_tg")}
    scaled = dict(params)
    for k, f in factors.items():
        scaled[k] = (params.get(k) or params["sigma"]) * f
    state = (phi, c, mu, scalars)
    pre = FixedStepProblem(scaled, cfg.initial_mu, state, dts, phi, every=every, device=device)
    phi_target = pre.final_phases(pre.theta_values())
    problem = FixedStepProblem(params, cfg.initial_mu, state, dts, phi_target, w_front=nx / cfg.dy ** 2,
                               every=every, device=device)
    return problem, factors, cfg
