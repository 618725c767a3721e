"""FP64 reference gradient for the adjoint tests: the fixed-size step
(propose_phi, then implicit_diffusion; pearl_serial.cpp:961-1094) restated in
torch and differentiated by autograd.

Written from the reference's equations, independently of the HIP adjoint:
sort-based simplex projection (:489-509), the same face weights and masks.
The diffusion output is mu' = mu_ref + x with x = A^{-1}(c - ic) and
c' = c - dt L x, the exact-solve form of :1082-1093.

Two solves for A x = b:
  "dense"  the operator as a dense n x n matrix and torch.linalg.solve
           (small grids only), differentiated by autograd through the solve;
  "cg"     the operator applied as a stencil and solved by Jacobi-PCG in
           torch; the solve is an autograd.Function differentiated
           implicitly: with lam = A^{-T} g (A is symmetric, so the same
           PCG), dL/db = lam and dL/dtheta_A = -lam^T (dA/dtheta_A) x, the
           last term by autograd of the stencil apply with x held fixed.
           Long runs are split into checkpointed segments
           (torch.utils.checkpoint), so the graph of one segment is held at
           a time; on a GPU this reaches the 512 x 256, 2,000-step case.
"""
import math

import torch
from torch.utils.checkpoint import checkpoint

# interface energy per phase pair (alpha-theta, alpha-gamma, theta-gamma),
# mobility coefficients, diffusivities.  The obstacle and gradient terms use
# sigma per pair; with equal pair energies they are the reference's terms.
THETA = ["sigma_at", "sigma_ag", "sigma_tg", "tau_at", "tau_ag", "tau_tg", "D0", "D1", "D2"]


def theta_of(pd):
    """THETA values of a parameter dict (a missing pair energy is sigma)."""
    return [float(pd.get(k) or pd["sigma"]) if k.startswith("sigma_") else float(pd[k]) for k in THETA]


def _shift(f, dy, dx):
    """f[..., y+dy, x+dx] with x periodic; rows outside the grid repeat the
    edge row (their faces are masked anyway)."""
    g = torch.roll(f, shifts=-dx, dims=-1) if dx else f
    if dy == 1:
        g = torch.cat([g[..., 1:, :], g[..., -1:, :]], dim=-2)
    elif dy == -1:
        g = torch.cat([g[..., :1, :], g[..., :-1, :]], dim=-2)
    return g


def _project(z):
    """Euclidean projection of z[3, ...] onto the probability simplex."""
    s, _ = torch.sort(z, dim=0, descending=True)
    cs = torch.cumsum(s, dim=0)
    j = torch.arange(1, 4, dtype=z.dtype, device=z.device).view(3, *([1] * (z.dim() - 1)))
    cand = (cs - 1.0) / j
    active = s > cand
    # largest active index
    idx = torch.where(active, j.expand_as(s) - 1, torch.full_like(s, -1)).max(dim=0).values.long()
    thr = torch.gather(cand, 0, idx.unsqueeze(0)).squeeze(0)
    return torch.clamp(z - thr, min=0.0)


def _faces(ny, nx, dx2, dy2, device):
    """(dy, dx, spacing^2, mask) of the west, east, south and north faces;
    the south face of row 0 and the north face of the last row are closed
    (no-flux), x is periodic."""
    yy = torch.arange(ny, device=device).view(ny, 1).expand(ny, nx)
    return [(0, -1, dx2, None), (0, 1, dx2, None), (-1, 0, dy2, yy > 0), (1, 0, dy2, yy < ny - 1)]


def _face_weights(K, dx2, dy2):
    ny, nx = K.shape
    out = []
    for (sy, sx, dd, mask) in _faces(ny, nx, dx2, dy2, K.device):
        w = 0.5 * (K + _shift(K, sy, sx)) / dd
        if mask is not None:
            w = torch.where(mask, w, torch.zeros_like(w))
        out.append((sy, sx, w))
    return out


def _laplace(K, x, dx2, dy2):
    """(L x)_i = sum over faces w_f (x_i - x_j)."""
    out = torch.zeros_like(x)
    for (sy, sx, w) in _face_weights(K, dx2, dy2):
        out = out + w * (x - _shift(x, sy, sx))
    return out


def _apply(chi, K, x, dt, dx2, dy2):
    return chi * x + dt * _laplace(K, x, dx2, dy2)


def _pcg(chi, K, b, dt, dx2, dy2, rtol=1e-15, max_it=5000):
    """Jacobi-PCG for A x = b; stops at |r| <= rtol |b| or when the residual
    has not decreased for 25 iterations (FP64 floor)."""
    diag = chi.clone()
    for (_, _, w) in _face_weights(K, dx2, dy2):
        diag = diag + dt * w
    x = torch.zeros_like(b)
    r = b.clone()
    z = r / diag
    p = z.clone()
    rz = (r * z).sum()
    bn = float(b.norm())
    if bn == 0.0:
        return x
    best, since = math.inf, 0
    for _ in range(max_it):
        Ap = _apply(chi, K, p, dt, dx2, dy2)
        a = rz / (p * Ap).sum()
        x = x + a * p
        r = r - a * Ap
        rn = float(r.norm())
        if rn <= rtol * bn:
            break
        if rn < best:
            best, since = rn, 0
        else:
            since += 1
            if since >= 25:
                break
        z = r / diag
        rz_new = (r * z).sum()
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x


class _CGSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, chi, K, b, dt, dx2, dy2):
        x = _pcg(chi, K, b, dt, dx2, dy2)
        ctx.save_for_backward(chi, K, x)
        ctx.geo = (dt, dx2, dy2)
        return x

    @staticmethod
    def backward(ctx, gx):
        chi, K, x = ctx.saved_tensors
        dt, dx2, dy2 = ctx.geo
        lam = _pcg(chi, K, gx, dt, dx2, dy2)
        with torch.enable_grad():
            chi_ = chi.detach().requires_grad_(True)
            K_ = K.detach().requires_grad_(True)
            y = _apply(chi_, K_, x.detach(), dt, dx2, dy2)
            gchi, gK = torch.autograd.grad(y, (chi_, K_), grad_outputs=-lam)
        return gchi, gK, lam, None, None, None


def _dense_solve(chi, K, b, dt, dx2, dy2):
    ny, nx = chi.shape
    n = nx * ny
    idx = torch.arange(n, device=chi.device).view(ny, nx)
    diag = chi.reshape(-1).clone()
    Amat = None
    entries = []
    for (sy, sx, w) in _face_weights(K, dx2, dy2):
        diag = diag + dt * w.reshape(-1)
        entries.append((idx.reshape(-1), _shift(idx, sy, sx).reshape(-1), -dt * w.reshape(-1)))
    Amat = torch.diag(diag)
    for r_, c_, v_ in entries:
        Amat = Amat.index_put((r_, c_), v_, accumulate=True)
    x = torch.linalg.solve(Amat, b.reshape(-1))
    # (L x)_i = sum_f w_f (x_i - x_j) = (A x - chi x)_i / dt
    Lx = (Amat @ x - chi.reshape(-1) * x) / dt
    return x.view(ny, nx), Lx.view(ny, nx)


def step(phi, c, mu, theta, pd, mu_ref, dt, solve="dense"):
    s_at, s_ag, s_tg, tau_at, tau_ag, tau_tg, D0, D1, D2 = theta
    sig = {(0, 1): s_at, (0, 2): s_ag, (1, 2): s_tg}
    eps, Vm, triple = pd["eps"], pd["Vm"], pd["triple"]
    A = [pd["A0"], pd["A1"], pd["A2"]]
    Bq = [pd["B0"], pd["B1"], pd["B2"]]
    Cq = [pd["C0"], pd["C1"], pd["C2"]]
    D = [D0, D1, D2]
    dx2, dy2 = pd["dx"] ** 2, pd["dy"] ** 2
    ny, nx = c.shape
    obs = {ab: 16 * v / (math.pi * math.pi * eps) for ab, v in sig.items()}
    tr = triple / eps
    fc = {ab: 2 * eps * v for ab, v in sig.items()}
    p = phi
    psi = [Cq[b] - 0.25 / A[b] * (mu - Bq[b]) ** 2 for b in range(3)]
    dpsi = [psi[b] - psi[2] for b in range(3)]
    F = []
    for a in range(3):
        a1, a2 = (a + 1) % 3, (a + 2) % 3
        dtr = 2.0 * p[a1] * p[a2]
        f = obs[tuple(sorted((a, a1)))] * p[a1] + obs[tuple(sorted((a, a2)))] * p[a2] + tr * p[a1] * p[a2]
        for b in range(3):
            dh = dtr + (6.0 * p[b] * (1.0 - p[b]) if b == a else 0.0)
            f = f + dpsi[b] * dh / Vm
        F.append(f)
    for (sy, sx, dd, mask) in _faces(ny, nx, dx2, dy2, c.device):
        q = _shift(phi, sy, sx)
        for a in range(3):
            for b in range(a + 1, 3):
                w = fc[(a, b)] * (p[a] * q[b] - q[a] * p[b]) / dd
                if mask is not None:
                    w = torch.where(mask, w, torch.zeros_like(w))
                F[a] = F[a] + w * q[b]
                F[b] = F[b] - w * q[a]
    mean = (F[0] + F[1] + F[2]) / 3
    w0, w1, w2 = p[0] * p[1], p[0] * p[2], p[1] * p[2]
    S = w0 + w1 + w2
    tmin = torch.minimum(torch.minimum(tau_at, tau_ag), tau_tg)
    tloc = torch.where(S > 1e-30, (w0 * tau_at + w1 * tau_ag + w2 * tau_tg) / torch.where(S > 1e-30, S, 1.0), tmin)
    s = dt / (tloc * eps)
    z = torch.stack([p[a] - s * (F[a] - mean) for a in range(3)])
    pt = _project(z)
    # diffusion on the trial phases
    h = [pt[q] * pt[q] * (3.0 - 2.0 * pt[q]) + 2.0 * pt[0] * pt[1] * pt[2] for q in range(3)]
    chi = sum(h[q] / (2.0 * A[q]) for q in range(3))
    c0 = -sum(h[q] * Bq[q] / (2.0 * A[q]) for q in range(3))
    K = sum(pt[q] * D[q] / (2.0 * A[q]) for q in range(3))
    ic = chi * mu_ref + c0
    b = c - ic
    if solve == "dense":
        x, Lx = _dense_solve(chi, K, b, dt, dx2, dy2)
    else:
        x = _CGSolve.apply(chi, K, b, dt, dx2, dy2)
        Lx = _laplace(K, x, dx2, dy2)
    c_new = c - dt * Lx
    mu_new = mu_ref + x
    return pt, c_new, mu_new


def loss_and_grad(pd, mu_ref, phi0, c0, mu0, dts, phi_target, w_front=0.0, front_target=0.0, solve="dense",
                  device="cpu", every=0):
    """Loss after the fixed steps and its autograd gradient w.r.t. THETA.

    every > 0 runs the steps in checkpointed segments of that many steps
    (the graph of one segment is rebuilt during the backward pass)."""
    dev = torch.device(device)
    th = torch.tensor(theta_of(pd), dtype=torch.float64, device=dev, requires_grad=True)
    phi = torch.tensor(phi0, dtype=torch.float64, device=dev)
    c = torch.tensor(c0, dtype=torch.float64, device=dev)
    mu = torch.tensor(mu0, dtype=torch.float64, device=dev)
    dts = [float(d) for d in dts]

    def run(phi, c, mu, theta, seg):
        for dt in seg:
            phi, c, mu = step(phi, c, mu, list(theta), pd, mu_ref, dt, solve)
        return phi, c, mu

    if every > 0:
        for k in range(0, len(dts), every):
            seg = dts[k:k + every]
            phi, c, mu = checkpoint(lambda a, b, m, t, seg=seg: run(a, b, m, t, seg), phi, c, mu, th,
                                    use_reentrant=False)
    else:
        phi, c, mu = run(phi, c, mu, th, dts)
    tgt = torch.tensor(phi_target, dtype=torch.float64, device=dev)
    ny, nx = c.shape
    H = (1.0 - phi[2]).sum() * pd["dy"] / nx
    L = 0.5 * ((phi - tgt) ** 2).sum() + 0.5 * w_front * (H - front_target) ** 2
    L.backward()
    return (L.item(), th.grad.cpu().numpy().copy(), phi.detach().cpu().numpy(), c.detach().cpu().numpy(),
            mu.detach().cpu().numpy())
