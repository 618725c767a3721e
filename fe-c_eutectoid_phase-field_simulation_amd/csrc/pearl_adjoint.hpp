// Discrete adjoint of the fixed-size step (propose_phi, then the implicit
// carbon diffusion; pearl_serial.cpp:961-1094) for one domain in FP64.
//
// Forward map of step n (state (phi, c, mu) -> (phi', c', mu')):
//   z   = phi - s(phi) * (F(phi, mu) - mean F),  s = dt / (tau_loc(phi) eps)
//   phi' = P(z)                                   (projection onto the simplex)
//   x   = A(phi')^{-1} (c - ic(phi')),           A = diag(chi(phi')) + dt L(K(phi'))
//   c'  = c - dt L(K(phi')) x                     (flux form)
//   mu' = mu_ref + x                              (= mu_ref + (c' - ic)/chi at convergence)
// with (L(K) y)_i = sum_f w_f (y_i - y_j), w_f = (K_i + K_j)/(2 d_f^2), y-edge
// faces masked.  The reverse step takes the loss gradients with respect to
// (phi', c', mu') and returns those with respect to (phi, c, mu), and adds
// the parameter gradients (sigma, tau_at, tau_ag, tau_tg, D_alpha,
// D_theta, D_gamma) to accumulators:
//   gx  = gmu' - dt L gc'                         (L symmetric)
//   A lam = gx                                    (the same Jacobi PCG)
//   gc  = gc' + lam
//   gK_m = -dt sum_{f at m} (x_i - x_j)(Y_i - Y_j)/(2 d_f^2),  Y = gc' + lam
//   gchi = -lam x,  gic = -lam
//   gphi' += sum_p (gchi + gic (mu_ref - B_p)) chi_p dh_p/dphi' + gK D chi
//   gz  = J_P^T gphi'  (on the active set: gphi'_a - mean over the set)
//   u = -gz,  v = s (u - mean u)
//   gphi = gz + sum_a v_a dF_a/dphi + (u.G) ds/dphi + (neighbour terms)
//   gmu  = sum_a v_a dF_a/dmu,  gsigma_ab = sum_a v_a F_a^(sigma_ab)/sigma_ab,
//   gtau_k = (u.G) ds/dtau_k,   gD_p = sum_i gK_i phi'_p chi_p
// The capillary force of a cell depends on its neighbours' phases; their
// gradient terms are gathered in a second pass (each cell collects what its
// four neighbours' forces owe it), so no atomics are needed.
#pragma once

#include "pearl_kernels.hpp"

namespace pearl_gpu {

// Parameter-gradient accumulators: sigma_at, sigma_ag, sigma_tg (interface
// energy per phase pair), tau_at, tau_ag, tau_tg, D_alpha, D_theta, D_gamma.
constexpr int ADJ_THETA = 9;

struct Adj {
    const double* phi0;  // state n (3 planes)
    const double* mu0;
    const double* phi1;  // state n+1 = the accepted trial (3 planes)
    const double* mu1;
    const double* gphi1;  // loss gradients w.r.t. state n+1
    const double* gc1;
    const double* gmu1;
    double* gphi0;  // loss gradients w.r.t. state n (written)
    double* gc0;
    double* gmu0;
    double* gphit;  // w.r.t. the trial phases (3 planes)
    double* v;      // s (u - mean u) per cell (3 planes)
    double* theta;  // ADJ_THETA accumulators
    double dt;
};

// dh_p/dphi_a of the interpolation weights h (weights :102-107).
__device__ __forceinline__ double dh(const double (&p)[3], int hp, int a) {
    const double other = a == 0 ? p[1] * p[2] : (a == 1 ? p[0] * p[2] : p[0] * p[1]);
    return (hp == a ? 6.0 * p[a] * (1.0 - p[a]) : 0.0) + 2.0 * other;
}

// ---------------------------------------------------------------------------
// Diffusion adjoint, part 1: the adjoint right-hand side gx = gmu' - dt L gc'
// and the PCG set-up on the trial phases (chi, K, Jacobi diagonal, x = 0,
// r = gx, the first direction p = z = r/diag, stored where the forward's
// prepare stores it); sums rhs.rhs, r.z, r.r for the loop head.
// ---------------------------------------------------------------------------
struct KG {
    double k, g;
};

template <int TY>
__global__ __launch_bounds__(TPB) void k_adj_prepare(Bufs B, Adj A) {
    const KP P = B.kp[0];
    const int lb = blockIdx.x, n = P.n;
    const double dt = A.dt;
    double* __restrict__ chi = B.chi;
    double* __restrict__ K = B.K;
    double* __restrict__ rhs = B.rhs;
    double* __restrict__ x = B.x;
    double* __restrict__ diag = B.diag;
    double* __restrict__ r = B.r;
    double* __restrict__ z = B.S->pidx ? B.p0 : B.p1;  // the first direction p = z
    double v[3] = {0.0, 0.0, 0.0};
    const Col col = column<TY>(P, lb, P.ny);
    march<TY, KG>(
        P, col,
        [&](const int j) {
            double p[3];
            load_phi(A.phi1, n, j, p);
            return KG{mixture_K(P, p), A.gc1[j]};
        },
        [&](const int i, const int y, const KG& s, const KG& c, const KG& nn) {
            double p[3], qw[3], qe[3];
            load_phi(A.phi1, n, i, p);
            load_phi(A.phi1, n, i + col.dw, qw);
            load_phi(A.phi1, n, i + col.de, qe);
            Mix m;
            (void)mixture(P, p, m);
            double wf[4];
            face_weights(P, y, c.k, mixture_K(P, qw), mixture_K(P, qe), s.k, nn.k, wf);
            const double gw = A.gc1[i + col.dw], ge = A.gc1[i + col.de];
            const double lg = wf[0] * (c.g - gw) + wf[1] * (c.g - ge) + wf[2] * (c.g - s.g) + wf[3] * (c.g - nn.g);
            const double b = A.gmu1[i] - dt * lg;
            const double d = m.s + dt * (wf[0] + wf[1] + wf[2] + wf[3]);
            chi[i] = m.s;
            K[i] = c.k;
            diag[i] = d;
            rhs[i] = b;
            x[i] = 0.0;
            r[i] = b;
            const double zz = b / d;
            z[i] = zz;
            v[0] += b * b;
            v[1] += b * zz;
            v[2] += b * b;
        });
    block_partial<3, 0u>(v, B.part + P.poff, P.nb, lb);
}

__global__ __launch_bounds__(TPB) void k_adj_prepare_fin(Bufs B) {
    const KP P = B.kp[0];
    DevScal* S = B.S;
    double v[3];
    if (!final_reduce<3, 0u>(v, B.part + P.poff, P.nb)) return;
    S->rhs2 = v[0];
    S->rz = v[1];
    S->rr = v[2];
    S->tol = P.cg_atol + P.cg_rtol * __builtin_sqrt(v[0]);
    S->restart = 1;
    S->pad = 0;
    S->x_mu = 0;
    S->x_lag = 0;
    cg_head(P, S, 0);
}

// ---------------------------------------------------------------------------
// Diffusion adjoint, part 2: lam = B.x (the solution); gc = gc' + lam; the
// gradients of chi, ic and K mapped onto the trial phases (added to gphi');
// gD partial sums.
// ---------------------------------------------------------------------------
struct XY {
    double x, y;
};

template <int TY>
__global__ __launch_bounds__(TPB) void k_adj_diffgrad(Bufs B, Adj A) {
    const KP P = B.kp[0];
    const int lb = blockIdx.x, n = P.n;
    const double dt = A.dt;
    const double* __restrict__ lam = B.x;
    double v[3] = {0.0, 0.0, 0.0};
    const Col col = column<TY>(P, lb, P.ny);
    auto at = [&](const int j) { return XY{A.mu1[j] - P.mu_ref, A.gc1[j] + lam[j]}; };
    march<TY, XY>(
        P, col, at, [&](const int i, const int y, const XY& s, const XY& c, const XY& nn) {
            const XY w = at(i + col.dw), e = at(i + col.de);
            const double hx = 0.5 * P.rdx2, hy = 0.5 * P.rdy2;
            double gK = (c.x - w.x) * (c.y - w.y) * hx + (c.x - e.x) * (c.y - e.y) * hx;
            if (y != 0) gK += (c.x - s.x) * (c.y - s.y) * hy;
            if (y != P.ny - 1) gK += (c.x - nn.x) * (c.y - nn.y) * hy;
            gK *= -dt;
            const double l = lam[i];
            A.gc0[i] = c.y;
            const double gchi = -l * c.x, gic = -l;
            double p[3];
            load_phi(A.phi1, n, i, p);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                double g = A.gphi1[a * n + i] + gK * P.D[a] * P.chi[a];
#pragma unroll
                for (int q = 0; q < 3; ++q) g += (gchi + gic * (P.mu_ref - P.B[q])) * P.chi[q] * dh(p, q, a);
                A.gphit[a * n + i] = g;
                v[a] += gK * p[a] * P.chi[a];
            }
        });
    block_partial<3, 0u>(v, B.part + P.poff, P.nb, lb);
}

__global__ __launch_bounds__(TPB) void k_adj_diffgrad_fin(Bufs B, Adj A) {
    const KP P = B.kp[0];
    double v[3];
    if (!final_reduce<3, 0u>(v, B.part + P.poff, P.nb)) return;
    A.theta[6] += v[0];
    A.theta[7] += v[1];
    A.theta[8] += v[2];
}

// ---------------------------------------------------------------------------
// Propose adjoint, pass A (per cell, own terms): recompute z and the active
// set exactly as the forward did, then gz, v, the local gradient of phi,
// gmu and the sigma / tau partial sums.
// ---------------------------------------------------------------------------
template <int TY>
__global__ __launch_bounds__(TPB) void k_adj_propose_a(Bufs B, Adj A) {
    const KP P = B.kp[0];
    const int lb = blockIdx.x, n = P.n;
    const double dt = A.dt;
    // sigma per pair (01, 02, 12), then tau_at, tau_ag, tau_tg
    double th[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const Col col = column<TY>(P, lb, P.ny);
    march<TY, Phases>(
        P, col,
        [&](const int j) {
            Phases q;
            load_phi(A.phi0, n, j, q.p);
            return q;
        },
        [&](const int i, const int y, const Phases& ps, const Phases& pc, const Phases& pn) {
            const double(&p)[3] = pc.p;
            double qw[3], qe[3];
            load_phi(A.phi0, n, i + col.dw, qw);
            load_phi(A.phi0, n, i + col.de, qe);
            const double m = A.mu0[i];
            double force[3], mean, scale, z[3];
            double north[3];
            DivConst div;
            propose_z(P, y, p, qw, qe, ps.p, pn.p, m, dt, force, mean, scale, z, nullptr, north, div);
            double gt[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) gt[a] = A.gphit[a * n + i];
            const bool finite = __builtin_isfinite(z[0]) && __builtin_isfinite(z[1]) && __builtin_isfinite(z[2]);
            double zp[3];
            if (!finite || !project_simplex(z, P.r3, zp)) {
                // the forward kept phi (non-finite trial): identity
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    A.gphi0[a * n + i] = gt[a];
                    A.v[a * n + i] = 0.0;
                }
                A.gmu0[i] = 0.0;
                return;
            }
            // J_P^T on the active set
            double gsum = 0.0;
            int na = 0;
#pragma unroll
            for (int a = 0; a < 3; ++a)
                if (zp[a] > 0.0) {
                    gsum += gt[a];
                    ++na;
                }
            double gz[3], u[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                gz[a] = zp[a] > 0.0 ? gt[a] - gsum / na : 0.0;
                u[a] = -gz[a];
            }
            const double ubar = (u[0] + u[1] + u[2]) / 3.0;
            double vv[3], G[3], UG = 0.0;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                vv[a] = scale * (u[a] - ubar);
                G[a] = force[a] - mean;
                UG += u[a] * G[a];
                A.v[a * n + i] = vv[a];
            }
            double g[3] = {gz[0], gz[1], gz[2]};
            // obstacle sum_{a<b} obs_ab (v_a p_b + v_b p_a) and triple tr p_a1 p_a2
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = a + 1; b < 3; ++b) {
                    g[a] += P.obs_p[pair_of(a, b)] * vv[b];
                    g[b] += P.obs_p[pair_of(a, b)] * vv[a];
                }
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int a1 = (a + 1) % 3, a2 = (a + 2) % 3;
                g[a1] += vv[a] * P.tr * p[a2];
                g[a2] += vv[a] * P.tr * p[a1];
            }
            // chemical sum_b dpsi_b dh_ab / Vm
            const double psi0 = psi(P, 2, m);
            double dpsi[3], ddpsi[3];
            const double dpsi0_dmu = 2.0 * P.qpsi[2] * (m - P.B[2]);
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                dpsi[b] = psi(P, b, m) - psi0;
                ddpsi[b] = 2.0 * P.qpsi[b] * (m - P.B[b]) - dpsi0_dmu;
            }
            const double Dsum = (dpsi[0] + dpsi[1] + dpsi[2]) * P.rVm;
            double gmu = 0.0;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int a1 = (a + 1) % 3, a2 = (a + 2) % 3;
                g[a1] += vv[a] * Dsum * 2.0 * p[a2];
                g[a2] += vv[a] * Dsum * 2.0 * p[a1];
                g[a] += vv[a] * dpsi[a] * 6.0 * (1.0 - 2.0 * p[a]) * P.rVm;
                const double dtr = 2.0 * p[a1] * p[a2];
                double dF = 0.0;
#pragma unroll
                for (int b = 0; b < 3; ++b) dF += ddpsi[b] * (dtr + (b == a ? 6.0 * p[b] * (1.0 - p[b]) : 0.0));
                gmu += vv[a] * dF * P.rVm;
            }
            // capillary, own-phase part.  Every sigma_ab-proportional force
            // term, contracted with v, goes to the pair's accumulator
            // (v . dF/dsigma_ab = that sum / sigma_ab).
            double vsig[3];
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = a + 1; b < 3; ++b) vsig[pair_of(a, b)] = P.obs_p[pair_of(a, b)] * (vv[a] * p[b] + vv[b] * p[a]);
#pragma unroll
            for (int dir = 0; dir < 4; ++dir) {
                if ((dir == 2 && y == 0) || (dir == 3 && y == P.ny - 1)) continue;
                const double(&q)[3] = dir == 0 ? qw : (dir == 1 ? qe : (dir == 2 ? ps.p : pn.p));
                const double rd = dir < 2 ? P.rdx2 : P.rdy2;
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int b = a + 1; b < 3; ++b) {
                        const double fc = P.fc_p[pair_of(a, b)] * rd;
                        const double om = vv[a] * q[b] - vv[b] * q[a];
                        g[a] += om * fc * q[b];
                        g[b] -= om * fc * q[a];
                        const double wv = fc * (p[a] * q[b] - q[a] * p[b]);
                        vsig[pair_of(a, b)] += wv * om;
                    }
            }
            // the step scale s = dt/(tau_loc eps)
            const double w0 = p[0] * p[1], w1 = p[0] * p[2], w2 = p[1] * p[2];
            const double S = w0 + w1 + w2;
            const double tl = S > 1e-30 ? (w0 * P.tau[0] + w1 * P.tau[1] + w2 * P.tau[2]) / S : P.tau_min;
            const double ds_dtl = -scale / tl;
            if (S > 1e-30) {
                const double dN[3] = {p[1] * P.tau[0] + p[2] * P.tau[1], p[0] * P.tau[0] + p[2] * P.tau[2],
                                      p[0] * P.tau[1] + p[1] * P.tau[2]};
                const double dS[3] = {p[1] + p[2], p[0] + p[2], p[0] + p[1]};
#pragma unroll
                for (int c = 0; c < 3; ++c) g[c] += UG * ds_dtl * (dN[c] - tl * dS[c]) / S;
                th[3] += UG * ds_dtl * w0 / S;
                th[4] += UG * ds_dtl * w1 / S;
                th[5] += UG * ds_dtl * w2 / S;
            } else {
                const int kmin = (P.tau[0] <= P.tau[1] && P.tau[0] <= P.tau[2]) ? 0 : (P.tau[1] <= P.tau[2] ? 1 : 2);
                th[3 + kmin] += UG * ds_dtl;
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) th[k] += vsig[k] / P.sig_p[k];
#pragma unroll
            for (int a = 0; a < 3; ++a) A.gphi0[a * n + i] = g[a];
            A.gmu0[i] = gmu;
        });
    block_partial<6, 0u>(th, B.part + P.poff, P.nb, lb);
}

__global__ __launch_bounds__(TPB) void k_adj_propose_a_fin(Bufs B, Adj A) {
    const KP P = B.kp[0];
    double v[6];
    if (!final_reduce<6, 0u>(v, B.part + P.poff, P.nb)) return;
    for (int q = 0; q < 6; ++q) A.theta[q] += v[q];
}

// Propose adjoint, pass B: each cell gathers what the capillary forces of
// its neighbours owe it.  Neighbour i with phases p and v(i) coupled to this
// cell (phases q) over a face of spacing d: for every pair a < b, with
// w = 2 eps sigma_ab (p_a q_b - q_a p_b)/d^2 and om = v_a q_b - v_b q_a,
//   g_q[b] += om * 2 eps sigma_ab p_a/d^2 + v_a w
//   g_q[a] -= om * 2 eps sigma_ab p_b/d^2 + v_b w
template <int TY>
__global__ __launch_bounds__(TPB) void k_adj_propose_b(Bufs B, Adj A) {
    const KP P = B.kp[0];
    const int lb = blockIdx.x, n = P.n;
    const Col col = column<TY>(P, lb, P.ny);
    for_column<TY>(P, col, [&](const int j, const int y) {
        double q[3], g[3];
        load_phi(A.phi0, n, j, q);
#pragma unroll
        for (int a = 0; a < 3; ++a) g[a] = A.gphi0[a * n + j];
#pragma unroll
        for (int dir = 0; dir < 4; ++dir) {
            if ((dir == 2 && y == 0) || (dir == 3 && y == P.ny - 1)) continue;
            const int i = dir == 0 ? j + col.dw : (dir == 1 ? j + col.de : (dir == 2 ? j - P.nx : j + P.nx));
            double p[3], vv[3];
            load_phi(A.phi0, n, i, p);
            load_phi(A.v, n, i, vv);
            const double rd = dir < 2 ? P.rdx2 : P.rdy2;
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = a + 1; b < 3; ++b) {
                    const double fc = P.fc_p[pair_of(a, b)] * rd;
                    const double w = fc * (p[a] * q[b] - q[a] * p[b]);
                    const double om = vv[a] * q[b] - vv[b] * q[a];
                    g[b] += om * fc * p[a] + vv[a] * w;
                    g[a] -= om * fc * p[b] + vv[b] * w;
                }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) A.gphi0[a * n + j] = g[a];
    });
}

}  // namespace pearl_gpu
