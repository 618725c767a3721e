"""Interface energy per phase pair (pearl_params.pair_sigma), the extension
the parameter-recovery case differentiates.

1. Pair energies given explicitly and all equal to sigma: the step is the
   reference's, bit for bit (advance over several steps).
2. Three different pair energies: propose_phi equals the torch restatement
   (tests/adjoint_reference.py, which writes the obstacle and gradient terms
   with sigma per pair) to 1e-14.
3. The free energy (reduced_energy, pearl_serial.cpp:793-821) is linear in
   the pair energies: E(pairs) - E(sigma) = dx dy sum_ab (sigma_ab - sigma)
   I_ab, with I_ab = 16/(pi^2 eps) sum_cells phi_a phi_b + eps sum_faces
   q_ab^2 (east faces periodic, north faces below the top row,
   q_ab = (phi_a n_b - n_a phi_b)/d), evaluated here in numpy."""
import math
import sys

import numpy as np
import pytest
import torch

from conftest import load_golden

pytestmark = pytest.mark.gpu
PAIRS = {"sigma_at": 0.55, "sigma_ag": 0.42, "sigma_tg": 0.61}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(pkg, g, pd):
    s = pkg.Solver(pd, float(g["sc0"][9]))
    s.set_state(g["phi0"], g["c0"], g["mu0"], g["sc0"][:9])
    return s


def test_equal_pairs_are_the_reference_step(pkg):
    g = load_golden("small_48x40")
    pd = dict(g["pd"])
    eq = dict(pd, sigma_at=pd["sigma"], sigma_ag=pd["sigma"], sigma_tg=pd["sigma"])
    with make(pkg, g, pd) as a, make(pkg, g, eq) as b:
        a.run(1e9, 8)
        b.run(1e9, 8)
        for x, y in zip(a.get_state()[:3], b.get_state()[:3]):
            assert np.array_equal(x, y)


def test_pair_propose_matches_restatement(pkg):
    import adjoint_reference as R
    g = load_golden("small_48x40")
    pd = dict(g["pd"], **PAIRS)
    dt = float(g["dt_probe"])
    with make(pkg, g, pd) as s:
        _, phit = s.propose_phi(dt)
    th = [torch.tensor(v, dtype=torch.float64) for v in R.theta_of(pd)]
    pt, _, _ = R.step(torch.tensor(g["phi0"]), torch.tensor(g["c0"]), torch.tensor(g["mu0"]), th, pd,
                      float(g["sc0"][9]), dt)
    diff = float(np.max(np.abs(pt.numpy() - phit)))
    say("pair propose vs restatement", diff)
    assert diff <= 1e-14
    # and it differs from the single-sigma step, so the pair energies are used
    with make(pkg, g, dict(g["pd"])) as s:
        _, phit1 = s.propose_phi(dt)
    assert float(np.max(np.abs(phit1 - phit))) > 1e-6


def interface_integrals(phi, pd):
    """I_ab per pair (01, 02, 12) per unit sigma, as a sum over cells."""
    eps, dx, dy = pd["eps"], pd["dx"], pd["dy"]
    obs1 = 16.0 / (math.pi * math.pi * eps)
    east = np.roll(phi, -1, axis=2)
    north = phi[:, 1:, :]
    out = []
    for a, b in ((0, 1), (0, 2), (1, 2)):
        bulk = obs1 * math.fsum((phi[a] * phi[b]).ravel())
        qe = (phi[a] * east[b] - east[a] * phi[b]) / dx
        qn = (phi[a, :-1] * north[b] - north[a] * phi[b, :-1]) / dy
        out.append(bulk + eps * (math.fsum((qe * qe).ravel()) + math.fsum((qn * qn).ravel())))
    return np.array(out)


def test_pair_energy_is_linear_in_pair_energies(pkg):
    g = load_golden("colony_256x128")
    pd = dict(g["pd"])
    with make(pkg, g, pd) as s:
        e0 = s.reduced_energy()
    with make(pkg, g, dict(pd, **PAIRS)) as s:
        e1 = s.reduced_energy()
    I = interface_integrals(g["phi0"], pd)
    ds = np.array([PAIRS["sigma_at"], PAIRS["sigma_ag"], PAIRS["sigma_tg"]]) - pd["sigma"]
    want = pd["dx"] * pd["dy"] * float(np.dot(ds, I))
    say("energy difference", e1 - e0, "numpy", want)
    assert abs((e1 - e0) - want) <= 1e-9 * abs(want)
