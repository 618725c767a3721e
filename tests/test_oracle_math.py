"""Mathematical pins of the oracle that do not need the reference build:
closed forms, invariants and brute force on tiny inputs."""
import itertools
import math

import numpy as np
import pytest

from conftest import load_golden


def tiny_params(O, nx=6, ny=5, **kw):
    g = load_golden("small_48x40")
    pd = dict(g["pd"])
    pd["nx"], pd["ny"] = nx, ny
    pd.update(kw)
    return O.params_from_dict(pd), pd


def brute_project(z):
    """Euclidean projection onto the probability simplex by enumerating the
    active sets (KKT), independent of the sort-based formula."""
    best, bd = None, np.inf
    for mask in itertools.product([0, 1], repeat=3):
        idx = [i for i in range(3) if mask[i]]
        if not idx:
            continue
        lam = (sum(z[i] for i in idx) - 1.0) / len(idx)
        x = np.zeros(3)
        for i in idx:
            x[i] = z[i] - lam
        if (x >= -1e-15).all():
            d = np.sum((x - z) ** 2)
            if d < bd:
                best, bd = x, d
    return best


def test_project_simplex_matches_bruteforce(O):
    rng = np.random.default_rng(1)
    for _ in range(2000):
        z = rng.normal(0.3, 0.6, 3)
        x = brute_project(z)
        rc, y = O.project_simplex(z)
        assert rc == 0
        assert np.allclose(x, y, atol=1e-12)
        assert abs(y.sum() - 1) < 1e-12 and (y >= 0).all()
    # points on the simplex are fixed points
    for _ in range(200):
        z = rng.dirichlet([1, 1, 1])
        rc, y = O.project_simplex(z)
        assert np.allclose(y, z, atol=1e-15)


def test_blocked_sum_tree(O):
    """The blocked order (tiles of 256 columns by k rows, one column per
    thread) on small arrays equals an explicit restatement of the tree, and
    stays within n*eps of the exact sum."""
    rng = np.random.default_rng(2)
    for nx, ny, k in [(1, 1, 1), (255, 1, 1), (256, 1, 1), (300, 7, 2), (48, 40, 4), (600, 37, 8),
                      (70, 1000, 1), (513, 9, 16)]:
        n = nx * ny
        v = rng.normal(size=n) * 10.0 ** rng.integers(-3, 3, size=n)
        P, _ = tiny_params(O, nx=nx, ny=ny)
        P.red_mode, P.red_k = O.RED_BLOCKED, k
        got = O.osum(P, v)
        # explicit restatement
        tiles_x = -(-nx // 256)
        tiles_y = -(-ny // k)
        parts = []
        for b in range(tiles_x * tiles_y):
            tx, y0 = b % tiles_x, (b // tiles_x) * k
            acc = []
            for t in range(256):
                x = tx * 256 + t
                s = 0.0
                if x < nx:
                    for y in range(y0, min(y0 + k, ny)):
                        s += float(v[y * nx + x])
                acc.append(s)
            parts.append(tree(acc))
        acc = []
        for t in range(256):
            s = 0.0
            for j in range(t, len(parts), 256):
                s += parts[j]
            acc.append(s)
        assert got == tree(acc)
        exact = math.fsum(v.tolist())
        assert abs(got - exact) <= n * 2.3e-16 * np.abs(v).sum() + 1e-300


def tree(a):
    a = list(a)
    w = []
    for wv in range(4):
        v = a[64 * wv:64 * wv + 64]
        s = 32
        while s >= 1:
            for l in range(s):
                v[l] = v[l] + v[l + s]
            s //= 2
        w.append(v[0])
    return (w[0] + w[1]) + (w[2] + w[3])


def test_diffusion_matches_dense_solve(O):
    """implicit_diffusion on a 6x5 grid equals the dense solution of the same
    SPD system (periodic x, zero-flux y, arithmetic face mobility), and it
    conserves carbon exactly in the flux form."""
    P, pd = tiny_params(O, nx=6, ny=5, cg_rtol=1e-14, cg_atol=1e-30)
    nx, ny = 6, 5
    rng = np.random.default_rng(3)
    phi = rng.dirichlet([1, 1, 1], size=(ny, nx)).transpose(2, 0, 1).copy()
    mu_ref = 31345.0
    mu = mu_ref + rng.normal(0, 5, (ny, nx))
    c0 = np.full((ny, nx), 0.03) + rng.normal(0, 1e-4, (ny, nx))
    dt = 1e-12
    d = O.implicit_diffusion(P, phi, c0, mu, dt, mu_ref)
    assert d["status"] == 0 and d["ok"] == 1
    A = [pd[f"A{i}"] for i in range(3)]
    B = [pd[f"B{i}"] for i in range(3)]
    D = [pd[f"D{i}"] for i in range(3)]
    h = [p * p * (3 - 2 * p) + 2 * phi[0] * phi[1] * phi[2] for p in phi]
    chi = sum(h[i] / (2 * A[i]) for i in range(3))
    cint = -sum(h[i] / (2 * A[i]) * B[i] for i in range(3))
    K = sum(phi[i] * D[i] / (2 * A[i]) for i in range(3))
    n = nx * ny
    M = np.zeros((n, n))
    idx = lambda x, y: y * nx + x
    for y in range(ny):
        for x in range(nx):
            i = idx(x, y)
            M[i, i] += chi[y, x]
            nb = [((x - 1) % nx, y, pd["dx"]), ((x + 1) % nx, y, pd["dx"])]
            if y > 0:
                nb.append((x, y - 1, pd["dy"]))
            if y < ny - 1:
                nb.append((x, y + 1, pd["dy"]))
            for xx, yy, h_ in nb:
                w = 0.5 * (K[y, x] + K[yy, xx]) / h_ ** 2
                M[i, i] += dt * w
                M[i, idx(xx, yy)] -= dt * w
    rhs = (c0 - (chi * mu_ref + cint)).ravel()
    u = np.linalg.solve(M, rhs)
    assert np.allclose(d["mu"].ravel() - mu_ref, u, rtol=1e-9, atol=1e-7)
    assert abs(d["carbon"].sum() - c0.sum()) <= 1e-15 * n


def test_weights_partition_of_unity(O):
    rng = np.random.default_rng(4)
    phi = rng.dirichlet([0.5, 0.5, 0.5], size=1000).T
    h = [p * p * (3 - 2 * p) + 2 * phi[0] * phi[1] * phi[2] for p in phi]
    assert np.allclose(sum(h), 1.0, atol=1e-14)


def test_chemical_force_invariance(O):
    """The proposal only depends on differences of the grand potentials:
    adding a constant to every C_p leaves phi' unchanged up to rounding."""
    g = load_golden("small_48x40")
    P = O.params_from_dict(g["pd"])
    ch0, phi0, _ = O.propose_phi(P, g["phi0"], g["mu0"], 1.5e-11)
    pd = dict(g["pd"])
    for i in range(3):
        pd[f"C{i}"] += 1000.0
    P2 = O.params_from_dict(pd)
    ch1, phi1, _ = O.propose_phi(P2, g["phi0"], g["mu0"], 1.5e-11)
    assert np.abs(phi1 - phi0).max() < 1e-9


def test_capillary_force_is_energy_gradient(O):
    """With the chemistry off (all psi equal), one explicit step lowers the
    interface energy (the force is the discrete energy gradient)."""
    g = load_golden("small_48x40")
    pd = dict(g["pd"])
    for i in range(3):
        pd[f"A{i}"], pd[f"B{i}"], pd[f"C{i}"] = 1e5, 0.0, 0.0
    P = O.params_from_dict(pd)
    mu = np.zeros_like(g["mu0"])
    c = np.zeros_like(g["c0"])
    pd2 = dict(pd)
    pd2["phase_only"] = 1
    P2 = O.params_from_dict(pd2)
    e0, _ = O.reduced_energy(P2, g["phi0"], c, 0.0)
    ch, phi1, _ = O.propose_phi(P, g["phi0"], mu, 1e-13)
    e1, _ = O.reduced_energy(P2, phi1, c, 0.0)
    assert e1 < e0
