"""Initial states for the lamellar-colony workloads (host side).

`sharp_labels` restates the FILLCUBE branch of filling_phase_at
(pearl_serial.cpp:1399-1405): one-hot phase labels, gamma background, later
commands overwrite earlier ones, inclusive integer bounds.

`diffuse_colony` builds a diffuse colony directly, without the capillarity
pre-smoothing pass: each cementite plate and the colony top get the
reference's compact obstacle profile (`profile`, :563-569), combined like
the reference's `lamella` initialization (:1767-1771):
  phi = ((1-th)*prod, th*prod, 1-prod).
Carbon starts at the common chemical potential mu0: c = chi*mu0 + c0 of the
local mixture (:1823-1826).
"""
from __future__ import annotations

import math

import numpy as np


def sharp_labels(nx: int, ny: int, fills) -> np.ndarray:
    """phi[3, ny, nx] from FILLCUBE commands [(method, params)]."""
    phi = np.zeros((3, ny, nx))
    phi[2] = 1.0
    for method, p in fills:
        if method != "FILLCUBE":
            raise NotImplementedError(method)
        sel = int(p[0])
        if sel == 2:
            continue  # the reference treats a gamma FILLCUBE as a no-op
        x0, y0, x1, y1 = int(p[1]), int(p[2]), int(p[4]), int(p[5])
        phi[:, y0:y1 + 1, x0:x1 + 1] = 0.0
        phi[sel, y0:y1 + 1, x0:x1 + 1] = 1.0
    return phi


def profile(distance: np.ndarray, eps: float) -> np.ndarray:
    """Compact obstacle profile (:563-569): 1 on the negative side."""
    pi = math.acos(-1.0)
    half = pi * pi * eps / 8
    out = 0.5 * (1 - np.sin(4 * distance / (pi * eps)))
    out = np.where(distance <= -half, 1.0, out)
    out = np.where(distance >= half, 0.0, out)
    return out


def colony_profiles(nx: int, ny: int, dx: float, dy: float, eps: float, plates, strip_rows: int):
    """The separable profiles of diffuse_colony: th[nx], the cementite
    profile of each column's signed distance to the nearest plate, and
    prod[ny], the colony profile of each row (the PROFILE_LAMELLA tables of
    Solver.init_profile)."""
    L = nx * dx
    xx = (np.arange(nx) + 0.5) * dx
    # signed distance to the nearest cementite plate (negative inside)
    dist = np.full(nx, np.inf)
    for x0, w in plates:
        a, b = x0 * dx, (x0 + w) * dx
        for shift in (-L, 0.0, L):
            inside = (xx >= a + shift) & (xx <= b + shift)
            d = np.where(inside, -np.minimum(xx - a - shift, b + shift - xx),
                         np.minimum(np.abs(xx - a - shift), np.abs(xx - b - shift)))
            dist = np.where(np.abs(d) < np.abs(dist), d, dist)
    yy = (np.arange(ny) + 0.5) * dy
    return profile(dist, eps), profile(yy - strip_rows * dy, eps)


def diffuse_colony(nx: int, ny: int, dx: float, dy: float, eps: float, plates, strip_rows: int) -> np.ndarray:
    """phi[3, ny, nx] for cementite plates [(x_start, width)] in cells, the
    colony occupying rows [0, strip_rows).  x is periodic."""
    th, prod = colony_profiles(nx, ny, dx, dy, eps, plates, strip_rows)
    th, prod = th[None, :], prod[:, None]
    phi = np.empty((3, ny, nx))
    phi[0] = (1 - th) * prod
    phi[1] = th * prod
    phi[2] = 1 - prod
    return phi


def plates_from_fills(fills):
    """Cementite plates [(x_start, width)] of a lamella FILLCUBE recipe."""
    out = []
    for method, p in fills:
        if isinstance(p, str):
            p = [float(v) for v in p.strip().strip(";").strip("{}").split(",")]
        if method == "FILLCUBE" and int(p[0]) == 1:
            out.append((int(p[1]), int(p[4]) - int(p[1]) + 1))
    return out


def carbon_at_mu(phi: np.ndarray, thermo, mu0: float) -> np.ndarray:
    """c = composition(mu0, mixture(phi)) per cell (:117-141)."""
    p0, p1, p2 = phi
    trip = 2.0 * p0 * p1 * p2
    h = [p * p * (3.0 - 2.0 * p) + trip for p in (p0, p1, p2)]
    s = np.zeros_like(p0)
    c0 = np.zeros_like(p0)
    for k in range(3):
        chi = 1.0 / (2.0 * thermo[k][0])
        s = s + h[k] * chi
        c0 = c0 - h[k] * chi * thermo[k][1]
    return s * mu0 + c0
