// Device-side pieces of the FP64 per-step solver: kernel parameter block,
// pointwise physics, the blocked reduction, and the kernels.
//
// Each per-cell expression keeps the reference's operation order and
// rounding (pearl_serial.cpp lines cited per function); the library is built
// with -ffp-contract=off so no multiply-add is fused unless the reference
// calls std::fma.  Loop-invariant quotients that the reference recomputes
// per call (1/(2A), -0.25/A, 16 sigma/(pi^2 eps), ...) are evaluated once on
// the host with the same expression, which yields the same bits.
#pragma once

#include <hip/hip_runtime.h>

namespace pearl_gpu {

constexpr int TPB = 256;  // threads per block of every cell kernel

// Kernel parameter block (by value).  All derived constants are formed on
// the host exactly as the reference forms them.
struct KP {
    int nx, ny, n;          // grid; n = nx*ny cells (< 2^31)
    int nb;                 // blocks of the cell kernels
    double dx, dy, dx2, dy2;  // dx2 = dx*dx (the reference's d.dx*d.dx)
    double A[3], B[3], C[3], D[3];
    double chi[3];          // 1.0/(2.0*A)        Parabola::chi   :84
    double qpsi[3];         // -0.25/A             Parabola::psi   :87
    double twoA[3];         // 2.0*A               Parabola::c_at_mu :83
    double obs, tr;         // 16*sigma/(pi*pi*eps), triple/eps   :964
    double face_coef;       // 2*eps*sigma                        :981
    double eps_sigma;       // eps*sigma                          :816
    double eps, Vm;
    double tau[3], tau_min;
    double psiref[3];       // psi_a(mu_ref) - psi_gamma(mu_ref), psiref[2]=0  :797
    double mu_ref;
    double max_phase_change, cg_rtol, cg_atol;  // controls used by device-side decisions
    int cg_max;
    int phase_only;
};

// Device-resident scalars written by the reduction finishers.
struct DevScal {
    double change;                        // propose_phi max |dphi|
    double rhs2, rz, rr, pAp, alpha, beta;  // PCG
    double tol, cg_res;                   // PCG tolerance (:1044) and info.residual
    double closure, valid, mass, energy;  // finalize + acceptance
    double e_valid, e_mass, e_energy;     // state reductions of the current state
    double front, gmax, removed;          // moving window
    double dmin_c, dmax_c, dmin_mu, dmax_mu, dsumerr, dmass, denergy;  // diagnostics
    int live;                             // attempt still running (phase increment passed)
    int cg_state;                         // CG_* below
    int cg_it;                            // info.it (:1048)
    int restart;                          // next direction is p = z (first iteration, restart)
    int pidx;                             // which p buffer holds the current direction
    int err;                              // ERRF_* bits
    unsigned counter;                     // last-block counter
    int pad;
    int kind[64];                         // per-launch record of the PCG kernels (profiling)
};

enum { ERRF_MIXTURE = 1, ERRF_PROJECTION = 2, ERRF_NONFINITE_CARBON = 4, ERRF_MIXTURE_STATE = 8 };

// PCG states (implicit_diffusion :1047-1080).  CG_ITER: take a step;
// CG_TRUE: recompute the true residual (:1053); CG_OK / CG_FAIL: done;
// CG_IDLE: no solve in this attempt.
enum { CG_IDLE = 0, CG_ITER = 1, CG_TRUE = 2, CG_OK = 3, CG_FAIL = 4 };
enum { KIND_NOOP = 0, KIND_ITER = 1, KIND_TRUE = 2 };

// The loop head of implicit_diffusion (:1047-1067) for iteration `it` with
// the current r.r: info.it, info.residual, then the convergence and
// iteration-limit branches.  Runs in one thread.
__device__ __forceinline__ void cg_head(const int cg_max, DevScal* S, int it) {
    const double res = __builtin_sqrt(S->rr);
    S->cg_it = it;
    S->cg_res = res;
    if (!__builtin_isfinite(res)) S->cg_state = CG_FAIL;
    else if (res <= S->tol) S->cg_state = CG_TRUE;
    else if (it == cg_max) S->cg_state = CG_FAIL;
    else S->cg_state = CG_ITER;
}

__device__ __forceinline__ double dmaxd(double a, double b) { return (a < b) ? b : a; }  // std::max
__device__ __forceinline__ double dmind(double a, double b) { return (b < a) ? b : a; }  // std::min

// ---------------------------------------------------------------------------
// Blocked reduction.  OPS packs 2 bits per quantity: 0 sum, 1 max, 2 min.
// Stage 1 (caller): each thread folds its cells into v[q] in cell order.
// Stage 2: 64-lane xor-halving tree, then (w0+w1)+(w2+w3) per block; block
// partials to part[q*nb+b].  The last block to arrive (atomic counter)
// reduces the partials in the same shape: thread t folds partials t, t+256,
// ... then the same tree.  Returns true in thread 0 of that block, with the
// final values in v.
// ---------------------------------------------------------------------------
template <unsigned OPS>
__device__ __forceinline__ int op_of(int q) { return (OPS >> (2 * q)) & 3; }

__device__ __forceinline__ double comb(int op, double a, double b) {
    return op == 0 ? a + b : (op == 1 ? dmaxd(a, b) : dmind(a, b));
}
__device__ __forceinline__ double ident(int op) {
    return op == 0 ? 0.0 : (op == 1 ? -__builtin_inf() : __builtin_inf());
}

template <int NQ, unsigned OPS>
__device__ __forceinline__ void wave_tree(double (&v)[NQ]) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double x = v[q];
#pragma unroll
        for (int s = 32; s >= 1; s >>= 1) x = comb(op_of<OPS>(q), x, __shfl_xor(x, s));
        v[q] = x;
    }
}

template <int NQ, unsigned OPS>
__device__ bool block_reduce(double (&v)[NQ], double* __restrict__ part, unsigned* counter, int nb) {
    __shared__ double sh[4][NQ];
    __shared__ int is_last;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    wave_tree<NQ, OPS>(v);
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) sh[w][q] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int op = op_of<OPS>(q);
            part[q * nb + blockIdx.x] = comb(op, comb(op, sh[0][q], sh[1][q]), comb(op, sh[2][q], sh[3][q]));
        }
        __threadfence();
        const unsigned prev = atomicAdd(counter, 1u);
        is_last = (prev == static_cast<unsigned>(nb) - 1u);
    }
    __syncthreads();
    if (!is_last) return false;
    __threadfence();
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int op = op_of<OPS>(q);
        double acc = ident(op);
        for (int b = threadIdx.x; b < nb; b += TPB)
            acc = comb(op, acc, __hip_atomic_load(&part[q * nb + b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        v[q] = acc;
    }
    wave_tree<NQ, OPS>(v);
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) sh[w][q] = v[q];
    }
    __syncthreads();
    if (threadIdx.x != 0) return false;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int op = op_of<OPS>(q);
        v[q] = comb(op, comb(op, sh[0][q], sh[1][q]), comb(op, sh[2][q], sh[3][q]));
    }
    *counter = 0u;
    return true;
}

// ---------------------------------------------------------------------------
// Pointwise physics (pearl_serial.cpp:76-145, :533-537, :491-509)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double psi(const KP& P, int p, double mu) {  // :85-89
    const double d = mu - P.B[p];
    return __builtin_fma(P.qpsi[p], d * d, P.C[p]);
}

__device__ __forceinline__ void weights(const double (&phi)[3], double (&h)[3]) {  // :102-107
    const double triple = 2.0 * phi[0] * phi[1] * phi[2];
#pragma unroll
    for (int p = 0; p < 3; ++p) h[p] = phi[p] * phi[p] * (3.0 - 2.0 * phi[p]) + triple;
}

struct Mix {
    double s, c0, K;
};

// mixture :125-138; returns false where the reference throws.
__device__ __forceinline__ bool mixture(const KP& P, const double (&phi)[3], Mix& m) {
    double h[3];
    weights(phi, h);
    m.s = 0.0;
    m.c0 = 0.0;
    m.K = 0.0;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const double chi = P.chi[p];
        m.s += h[p] * chi;
        m.c0 -= h[p] * chi * P.B[p];
        m.K += phi[p] * P.D[p] * chi;
    }
    return (m.s > 0.0) && __builtin_isfinite(m.s) && __builtin_isfinite(m.c0) && __builtin_isfinite(m.K) &&
           !(m.K < 0.0);
}

__device__ __forceinline__ double composition(double mu, const Mix& m) {  // :139-141
    return __builtin_fma(m.s, mu, m.c0);
}

__device__ __forceinline__ double local_tau(const KP& P, const double (&p)[3]) {  // :533-537
    const double w0 = p[0] * p[1], w1 = p[0] * p[2], w2 = p[1] * p[2];
    const double s = w0 + w1 + w2;
    return s > 1e-30 ? (w0 * P.tau[0] + w1 * P.tau[1] + w2 * P.tau[2]) / s : P.tau_min;
}

// project_simplex :491-509; false where the reference throws.
__device__ __forceinline__ bool project_simplex(const double (&z)[3], double (&out)[3]) {
    double s0 = z[0], s1 = z[1], s2 = z[2], t;
    if (s0 < s1) { t = s0; s0 = s1; s1 = t; }
    if (s1 < s2) { t = s1; s1 = s2; s2 = t; }
    if (s0 < s1) { t = s0; s0 = s1; s1 = t; }
    double sum = 0.0, threshold = 0.0;
    int active = 0;
    sum += s0;
    { const double cand = (sum - 1.0) / 1.0; if (s0 > cand) { active = 1; threshold = cand; } }
    sum += s1;
    { const double cand = (sum - 1.0) / 2.0; if (s1 > cand) { active = 2; threshold = cand; } }
    sum += s2;
    { const double cand = (sum - 1.0) / 3.0; if (s2 > cand) { active = 3; threshold = cand; } }
#pragma unroll
    for (int a = 0; a < 3; ++a) out[a] = dmaxd(0.0, z[a] - threshold);
    return active != 0;
}

// Cell coordinates and stencil neighbours: x periodic; at the y edges the
// neighbour index falls back to the cell itself (the reference's y ghosts
// are copies of the edge row, :692-693).
struct Nbr {
    int x, y, w, e, s, nn;
};
__device__ __forceinline__ Nbr neighbours(const KP& P, int i) {
    Nbr r;
    r.y = i / P.nx;
    r.x = i - r.y * P.nx;
    r.w = r.x == 0 ? i + P.nx - 1 : i - 1;
    r.e = r.x == P.nx - 1 ? i - P.nx + 1 : i + 1;
    r.s = r.y == 0 ? i : i - P.nx;
    r.nn = r.y == P.ny - 1 ? i : i + P.nx;
    return r;
}

__device__ __forceinline__ void load_phi(const double* __restrict__ phi, int n, int i, double (&p)[3]) {
    p[0] = phi[i];
    p[1] = phi[n + i];
    p[2] = phi[2 * n + i];
}

// weights_at :1022-1028
__device__ __forceinline__ void face_weights(const KP& P, const double* __restrict__ K, int i, const Nbr& b,
                                             double (&w)[4]) {
    const double k = K[i];
    w[0] = .5 * (k + K[b.w]) / P.dx2;
    w[1] = .5 * (k + K[b.e]) / P.dx2;
    w[2] = .5 * (k + K[b.s]) / P.dy2;
    w[3] = .5 * (k + K[b.nn]) / P.dy2;
    if (b.y == 0) w[2] = 0;
    if (b.y == P.ny - 1) w[3] = 0;
}

// apply :1029-1035
__device__ __forceinline__ double apply_cell(const KP& P, const double* __restrict__ in, double chi, double dt,
                                             int i, const Nbr& b, const double (&w)[4]) {
    const double v = in[i];
    return chi * v + dt * (w[0] * (v - in[b.w]) + w[1] * (v - in[b.e]) + w[2] * (v - in[b.s]) + w[3] * (v - in[b.nn]));
}

// Per-cell acceptance quantities of valid_state (:1098-1114) and the local
// energy of reduced_energy (:799-819): the cell term, then the east and
// north face terms in the reference's order.
__device__ __forceinline__ void accept_cell(const KP& P, const double* __restrict__ phi, int i, const Nbr& b,
                                            double c, double mu, bool& good, double& energy, bool& mix_ok) {
    const int n = P.n;
    double p[3];
    load_phi(phi, n, i, p);
    double sum = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        good = good && __builtin_isfinite(p[a]) && p[a] >= -1e-12 && p[a] <= 1 + 1e-12;
        sum += p[a];
    }
    good = good && __builtin_fabs(sum - 1) < 1e-10;
    good = good && __builtin_isfinite(mu) && __builtin_isfinite(c);
    if (!P.phase_only) {
        good = good && c >= -1e-10 && c <= 1 + 1e-10;
#pragma unroll
        for (int a = 0; a < 3; ++a)
            if (p[a] > .01) {
                const double cp = (mu - P.B[a]) / P.twoA[a];
                good = good && cp >= -1e-10 && cp <= 1 + 1e-10;
            }
    }
    double h[3];
    weights(p, h);
    double chem = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) chem += h[a] * P.psiref[a];
    if (!P.phase_only) {
        Mix m;
        if (!mixture(P, p, m)) mix_ok = false;
        const double dc = c - composition(P.mu_ref, m);
        chem += .5 * dc * dc / m.s;
    }
    double v = chem / P.Vm + P.obs * (p[0] * p[1] + p[0] * p[2] + p[1] * p[2]) + P.tr * p[0] * p[1] * p[2];
    {
        double q[3];
        load_phi(phi, n, b.e, q);
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int bb = a + 1; bb < 3; ++bb) {
                const double z = (p[a] * q[bb] - q[a] * p[bb]) / P.dx;
                v += P.eps_sigma * z * z;
            }
    }
    if (b.y != P.ny - 1) {
        double q[3];
        load_phi(phi, n, b.nn, q);
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int bb = a + 1; bb < 3; ++bb) {
                const double z = (p[a] * q[bb] - q[a] * p[bb]) / P.dy;
                v += P.eps_sigma * z * z;
            }
    }
    energy = v;
}

// ---------------------------------------------------------------------------
// Kernels.  KK = cells per thread; block b owns cells [b*256*KK, (b+1)*256*KK).
// ---------------------------------------------------------------------------
#define PEARL_CELL_LOOP(KK)                                                   \
    _Pragma("unroll") for (int j_ = 0; j_ < (KK); ++j_) {                     \
        const int i = (blockIdx.x * (KK) + j_) * TPB + threadIdx.x;          \
        if (i >= P.n) break;

#define PEARL_CELL_LOOP_END }

// P1: propose_phi :961-1002.
template <int KK>
__global__ __launch_bounds__(TPB) void k_propose(KP P, const double* __restrict__ phi, const double* __restrict__ mu,
                                                double* __restrict__ out, double dt, double* part, DevScal* S) {
    const int n = P.n;
    double v[1] = {0.0};
    bool proj_fail = false;
    PEARL_CELL_LOOP(KK)
    const Nbr nb = neighbours(P, i);
    double p[3], force[3];
    load_phi(phi, n, i, p);
    const double m = mu[i];
    const double psi0 = psi(P, 2, m);
    const double dtr[3] = {2.0 * p[1] * p[2], 2.0 * p[0] * p[2], 2.0 * p[0] * p[1]};
    double dpsi[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) dpsi[b] = psi(P, b, m) - psi0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int a1 = (a + 1) % 3, a2 = (a + 2) % 3;
        force[a] = P.obs * (p[a1] + p[a2]) + P.tr * p[a1] * p[a2];
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const double dh = dtr[a] + (b == a ? 6.0 * p[b] * (1.0 - p[b]) : 0.0);
            force[a] += dpsi[b] * dh / P.Vm;
        }
    }
#pragma unroll
    for (int dir = 0; dir < 4; ++dir) {
        if ((dir == 2 && nb.y == 0) || (dir == 3 && nb.y == P.ny - 1)) continue;
        const int j = dir == 0 ? nb.w : (dir == 1 ? nb.e : (dir == 2 ? nb.s : nb.nn));
        double q[3];
        load_phi(phi, n, j, q);
        const double dd = dir < 2 ? P.dx2 : P.dy2;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = a + 1; b < 3; ++b) {
                const double w = P.face_coef * (p[a] * q[b] - q[a] * p[b]) / dd;
                force[a] += w * q[b];
                force[b] -= w * q[a];
            }
    }
    const double mean = (force[0] + force[1] + force[2]) / 3, scale = dt / (local_tau(P, p) * P.eps);
    double z[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) z[a] = p[a] - scale * (force[a] - mean);
    if (!(__builtin_isfinite(z[0]) && __builtin_isfinite(z[1]) && __builtin_isfinite(z[2]))) {
        v[0] = __builtin_inf();
#pragma unroll
        for (int a = 0; a < 3; ++a) out[a * n + i] = p[a];
        continue;
    }
    double zp[3];
    if (!project_simplex(z, zp)) proj_fail = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) v[0] = dmaxd(v[0], __builtin_fabs(zp[a] - p[a]));
#pragma unroll
    for (int a = 0; a < 3; ++a) out[a * n + i] = zp[a];
    PEARL_CELL_LOOP_END
    if (proj_fail) atomicOr(&S->err, ERRF_PROJECTION);
    if (block_reduce<1, 1u>(v, part, &S->counter, P.nb)) {
        S->change = v[0];
        // :1131 -- the rest of the attempt runs only if the increment passes
        S->live = (__builtin_isfinite(v[0]) && v[0] <= P.max_phase_change) ? 1 : 0;
        S->cg_state = CG_IDLE;
    }
}

// P2: implicit_diffusion setup :1013-1020.
template <int KK>
__global__ __launch_bounds__(TPB) void k_setup(KP P, const double* __restrict__ phi_t, const double* __restrict__ c_old,
                                              const double* __restrict__ mu_old, double* __restrict__ chi,
                                              double* __restrict__ icpt, double* __restrict__ K,
                                              double* __restrict__ rhs, double* __restrict__ x, DevScal* S) {
    if (!S->live) return;
    bool bad = false;
    PEARL_CELL_LOOP(KK)
    double p[3];
    load_phi(phi_t, P.n, i, p);
    Mix m;
    if (!mixture(P, p, m)) bad = true;
    const double ic = composition(P.mu_ref, m);
    chi[i] = m.s;
    icpt[i] = ic;
    K[i] = m.K;
    rhs[i] = c_old[i] - ic;
    x[i] = mu_old[i] - P.mu_ref;
    PEARL_CELL_LOOP_END
    if (bad) atomicOr(&S->err, ERRF_MIXTURE);
}

// P3: initial residual, Jacobi diagonal, preconditioned residual
// (:1036-1045).  Sums rhs.rhs, r.z, r.r; the finisher forms the tolerance
// and evaluates the loop head for iteration 0.
template <int KK>
__global__ __launch_bounds__(TPB) void k_residual0(KP P, const double* __restrict__ chi, const double* __restrict__ K,
                                                  const double* __restrict__ rhs, const double* __restrict__ x,
                                                  double* __restrict__ diag, double* __restrict__ r,
                                                  double* __restrict__ z, double dt, double* part, DevScal* S) {
    if (!S->live) return;
    double v[3] = {0.0, 0.0, 0.0};
    PEARL_CELL_LOOP(KK)
    const Nbr nb = neighbours(P, i);
    double w[4];
    face_weights(P, K, i, nb, w);
    const double c = chi[i];
    const double ax = apply_cell(P, x, c, dt, i, nb, w);
    const double d = c + dt * (w[0] + w[1] + w[2] + w[3]);
    const double b = rhs[i];
    const double rr = b - ax;
    const double zz = rr / d;
    diag[i] = d;
    r[i] = rr;
    z[i] = zz;
    v[0] += b * b;
    v[1] += rr * zz;
    v[2] += rr * rr;
    PEARL_CELL_LOOP_END
    if (block_reduce<3, 0u>(v, part, &S->counter, P.nb)) {
        S->rhs2 = v[0];
        S->rz = v[1];
        S->rr = v[2];
        S->tol = P.cg_atol + P.cg_rtol * __builtin_sqrt(v[0]);
        S->restart = 1;
        cg_head(P.cg_max, S, 0);
    }
}

// PCG kernel A.  In state CG_ITER: the search direction of this iteration
// p = z + beta p_old (:1079; p = z after a (re)start :1042/:1062), kept in
// the other p buffer, then Ap = A p (:1068) and p.Ap; the finisher checks
// for breakdown (:1070) and forms the step length (:1071).  In state
// CG_TRUE: the true residual r = rhs - A x with z = r/diag (:1053-1062);
// the finisher accepts (:1056) or restarts from it (:1060-1065).
// `slot` indexes the per-launch record used by the profiler.
template <int KK>
__global__ __launch_bounds__(TPB) void k_cg_a(KP P, const double* __restrict__ chi, const double* __restrict__ K,
                                             const double* __restrict__ rhs, const double* __restrict__ x,
                                             const double* __restrict__ diag, double* __restrict__ r,
                                             double* __restrict__ z, double* __restrict__ p0,
                                             double* __restrict__ p1, double* __restrict__ Ap, double dt, int slot,
                                             double* part, DevScal* S) {
    const int state = S->cg_state;
    if (blockIdx.x == 0 && threadIdx.x == 0 && slot >= 0)
        S->kind[slot] = state == CG_ITER ? KIND_ITER : (state == CG_TRUE ? KIND_TRUE : KIND_NOOP);
    if (state == CG_ITER) {
        const int pi = S->pidx;
        const double* __restrict__ po = pi ? p1 : p0;
        double* __restrict__ pn = pi ? p0 : p1;
        const bool fresh = S->restart != 0;
        const double beta = S->beta;
        double v[1] = {0.0};
        PEARL_CELL_LOOP(KK)
        const Nbr nb = neighbours(P, i);
        double w[4];
        face_weights(P, K, i, nb, w);
        const double pc = fresh ? z[i] : z[i] + beta * po[i];
        const double pw = fresh ? z[nb.w] : z[nb.w] + beta * po[nb.w];
        const double pe = fresh ? z[nb.e] : z[nb.e] + beta * po[nb.e];
        const double ps = fresh ? z[nb.s] : z[nb.s] + beta * po[nb.s];
        const double pnn = fresh ? z[nb.nn] : z[nb.nn] + beta * po[nb.nn];
        const double a = chi[i] * pc + dt * (w[0] * (pc - pw) + w[1] * (pc - pe) + w[2] * (pc - ps) + w[3] * (pc - pnn));
        pn[i] = pc;
        Ap[i] = a;
        v[0] += pc * a;
        PEARL_CELL_LOOP_END
        if (block_reduce<1, 0u>(v, part, &S->counter, P.nb)) {
            const double pAp = v[0], rz = S->rz;
            S->pAp = pAp;
            if (!(pAp > 0) || !__builtin_isfinite(pAp) || !__builtin_isfinite(rz)) {
                S->cg_state = CG_FAIL;
            } else {
                S->alpha = rz / pAp;
                S->pidx = pi ^ 1;
            }
        }
    } else if (state == CG_TRUE) {
        double v[2] = {0.0, 0.0};
        PEARL_CELL_LOOP(KK)
        const Nbr nb = neighbours(P, i);
        double w[4];
        face_weights(P, K, i, nb, w);
        const double rr = rhs[i] - apply_cell(P, x, chi[i], dt, i, nb, w);
        const double zz = rr / diag[i];
        r[i] = rr;
        z[i] = zz;
        v[0] += rr * rr;
        v[1] += rr * zz;
        PEARL_CELL_LOOP_END
        if (block_reduce<2, 0u>(v, part, &S->counter, P.nb)) {
            S->rr = v[0];
            const double res = __builtin_sqrt(v[0]);
            S->cg_res = res;
            if (res <= S->tol * 1.05) {
                S->cg_state = CG_OK;
            } else {
                S->rz = v[1];
                S->restart = 1;
                cg_head(P.cg_max, S, S->cg_it + 1);
            }
        }
    }
}

// PCG kernel B (CG_ITER only): x += alpha p; r -= alpha Ap; z = r/diag; sums
// r.z and r.r; the finisher forms beta (:1077) and evaluates the loop head
// of the next iteration.
template <int KK>
__global__ __launch_bounds__(TPB) void k_cg_b(KP P, double* __restrict__ x, double* __restrict__ r,
                                             double* __restrict__ z, const double* __restrict__ p0,
                                             const double* __restrict__ p1, const double* __restrict__ Ap,
                                             const double* __restrict__ diag, int slot, double* part, DevScal* S) {
    const int state = S->cg_state;
    if (blockIdx.x == 0 && threadIdx.x == 0 && slot >= 0) S->kind[slot] = state == CG_ITER ? KIND_ITER : KIND_NOOP;
    if (state != CG_ITER) return;
    const double* __restrict__ p = S->pidx ? p1 : p0;
    const double alpha = S->alpha;
    double v[2] = {0.0, 0.0};
    PEARL_CELL_LOOP(KK)
    x[i] += alpha * p[i];
    const double rr = r[i] - alpha * Ap[i];
    const double zz = rr / diag[i];
    r[i] = rr;
    z[i] = zz;
    v[0] += rr * zz;
    v[1] += rr * rr;
    PEARL_CELL_LOOP_END
    if (block_reduce<2, 0u>(v, part, &S->counter, P.nb)) {
        const double next = v[0];
        S->beta = next / S->rz;
        S->rz = next;
        S->rr = v[1];
        S->restart = 0;
        cg_head(P.cg_max, S, S->cg_it + 1);
    }
}

// F + A: flux-form carbon update, closure, mu reconstruction (:1082-1093),
// fused with the acceptance sweep on the trial state: valid_state (:1096),
// total_carbon (:784) and reduced_energy (:793).
// Reductions: closure max, valid min, carbon sum, energy sum.
template <int KK>
__global__ __launch_bounds__(TPB) void k_finalize(KP P, const double* __restrict__ phi_t,
                                                 const double* __restrict__ chi, const double* __restrict__ icpt,
                                                 const double* __restrict__ K, const double* __restrict__ x,
                                                 const double* __restrict__ c_old, double* __restrict__ c_new,
                                                 double* __restrict__ mu_new, double dt, double* part, DevScal* S) {
    if (S->cg_state != CG_OK) return;
    double v[4] = {0.0, 1.0, 0.0, 0.0};
    bool good = true, mix_ok = true;
    PEARL_CELL_LOOP(KK)
    const Nbr nb = neighbours(P, i);
    double w[4];
    face_weights(P, K, i, nb, w);
    const double xi = x[i];
    const double change = dt * (w[0] * (x[nb.w] - xi) + w[1] * (x[nb.e] - xi) + w[2] * (x[nb.s] - xi) + w[3] * (x[nb.nn] - xi));
    const double cn = c_old[i] + change;
    const double ic = icpt[i], ch = chi[i];
    v[0] = dmaxd(v[0], __builtin_fabs(cn - (ic + ch * xi)));
    const double mn = P.mu_ref + (cn - ic) / ch;
    c_new[i] = cn;
    mu_new[i] = mn;
    double e;
    accept_cell(P, phi_t, i, nb, cn, mn, good, e, mix_ok);
    v[2] += cn;
    v[3] += e;
    PEARL_CELL_LOOP_END
    v[1] = good ? 1.0 : 0.0;
    if (!mix_ok) atomicOr(&S->err, ERRF_MIXTURE);
    if (block_reduce<4, (1u | (2u << 2))>(v, part, &S->counter, P.nb)) {
        S->closure = v[0];
        S->valid = v[1];
        S->mass = v[2];
        S->energy = v[3];
    }
}

// Phase-only acceptance (:1134-1136): carbon = composition(mu, mixture(phi')),
// mu unchanged; then valid_state and reduced_energy on the trial state.
template <int KK>
__global__ __launch_bounds__(TPB) void k_phase_only(KP P, const double* __restrict__ phi_t,
                                                   const double* __restrict__ mu_old, double* __restrict__ c_new,
                                                   double* __restrict__ mu_new, double* part, DevScal* S) {
    if (!S->live) return;
    double v[4] = {0.0, 1.0, 0.0, 0.0};
    bool good = true, mix_ok = true;
    PEARL_CELL_LOOP(KK)
    const Nbr nb = neighbours(P, i);
    double p[3];
    load_phi(phi_t, P.n, i, p);
    Mix m;
    if (!mixture(P, p, m)) mix_ok = false;
    const double mu = mu_old[i];
    const double cn = composition(mu, m);
    c_new[i] = cn;
    mu_new[i] = mu;
    double e;
    accept_cell(P, phi_t, i, nb, cn, mu, good, e, mix_ok);
    v[2] += cn;
    v[3] += e;
    PEARL_CELL_LOOP_END
    v[1] = good ? 1.0 : 0.0;
    if (!mix_ok) atomicOr(&S->err, ERRF_MIXTURE);
    if (block_reduce<4, (1u | (2u << 2))>(v, part, &S->counter, P.nb)) {
        S->closure = v[0];
        S->valid = v[1];
        S->mass = v[2];
        S->energy = v[3];
    }
}

// valid_state, total_carbon and reduced_energy of a state (e0, mass0 and the
// single-function entry points).
template <int KK>
__global__ __launch_bounds__(TPB) void k_state_reduce(KP P, const double* __restrict__ phi,
                                                     const double* __restrict__ c, const double* __restrict__ mu,
                                                     double* part, DevScal* S) {
    double v[3] = {1.0, 0.0, 0.0};
    bool good = true, mix_ok = true;
    PEARL_CELL_LOOP(KK)
    const Nbr nb = neighbours(P, i);
    double e;
    const double ci = c[i];
    accept_cell(P, phi, i, nb, ci, mu[i], good, e, mix_ok);
    v[1] += ci;
    v[2] += e;
    PEARL_CELL_LOOP_END
    v[0] = good ? 1.0 : 0.0;
    if (!mix_ok) atomicOr(&S->err, ERRF_MIXTURE_STATE);
    if (block_reduce<3, 2u>(v, part, &S->counter, P.nb)) {
        S->e_valid = v[0];
        S->e_mass = v[1];
        S->e_energy = v[2];
    }
}

// Snapshot reductions (:2267-2282): min/max of carbon and mu, max phase-sum
// error, carbon sum, energy.
template <int KK>
__global__ __launch_bounds__(TPB) void k_diagnostics(KP P, const double* __restrict__ phi,
                                                    const double* __restrict__ c, const double* __restrict__ mu,
                                                    double* part, DevScal* S) {
    double v[7] = {1e300, -1e300, 1e300, -1e300, 0.0, 0.0, 0.0};
    bool good = true, mix_ok = true;
    PEARL_CELL_LOOP(KK)
    const Nbr nb = neighbours(P, i);
    const double ci = c[i], mi = mu[i];
    v[0] = dmind(v[0], ci);
    v[1] = dmaxd(v[1], ci);
    v[2] = dmind(v[2], mi);
    v[3] = dmaxd(v[3], mi);
    double p[3];
    load_phi(phi, P.n, i, p);
    v[4] = dmaxd(v[4], __builtin_fabs(p[0] + p[1] + p[2] - 1));
    double e;
    accept_cell(P, phi, i, nb, ci, mi, good, e, mix_ok);
    v[5] += ci;
    v[6] += e;
    PEARL_CELL_LOOP_END
    if (!mix_ok) atomicOr(&S->err, ERRF_MIXTURE);
    constexpr unsigned OPS = 2u | (1u << 2) | (2u << 4) | (1u << 6) | (1u << 8);
    if (block_reduce<7, OPS>(v, part, &S->counter, P.nb)) {
        S->dmin_c = v[0];
        S->dmax_c = v[1];
        S->dmin_mu = v[2];
        S->dmax_mu = v[3];
        S->dsumerr = v[4];
        S->dmass = v[5];
        S->denergy = v[6];
    }
}

// Moving-window front (:2057-2062): max over interior faces of the gamma=1/2
// crossing height.
template <int KK>
__global__ __launch_bounds__(TPB) void k_front(KP P, const double* __restrict__ phi, double* part, DevScal* S) {
    const double* g = phi + 2 * P.n;
    double v[1] = {-1.0};
    PEARL_CELL_LOOP(KK)
    const int y = i / P.nx;
    if (y < P.ny - 1) {
        const double a = 1 - 2 * g[i], b = 1 - 2 * g[i + P.nx];
        if (a >= 0 && b < 0) v[0] = dmaxd(v[0], ((y + 1) - .5 + a / (a - b)) * P.dy);
    }
    PEARL_CELL_LOOP_END
    if (block_reduce<1, 1u>(v, part, &S->counter, P.nb)) S->front = v[0];
}

// Max gamma and removed carbon (c*dx)*dy over the first m cells (:2028, :2067).
// Launched with ceil(m/(256*KK)) blocks; nb is that count.
template <int KK>
__global__ __launch_bounds__(TPB) void k_strip(KP P, const double* __restrict__ phi, const double* __restrict__ c,
                                              int m, int nb, double* part, DevScal* S) {
    const double* g = phi + 2 * P.n;
    double v[2] = {0.0, 0.0};
    _Pragma("unroll") for (int j_ = 0; j_ < KK; ++j_) {
        const int i = (blockIdx.x * KK + j_) * TPB + threadIdx.x;
        if (i >= m) break;
        v[0] = dmaxd(v[0], g[i]);
        v[1] += c[i] * P.dx * P.dy;
    }
    if (block_reduce<2, 1u>(v, part, &S->counter, nb)) {
        S->gmax = v[0];
        S->removed = v[1];
    }
}

// shift_rows row move and gamma fill (:2029-2042) into the other buffer,
// then close_mu (:789-792) on the new fields.
template <int KK>
__global__ __launch_bounds__(TPB) void k_shift(KP P, const double* __restrict__ phi_in, const double* __restrict__ c_in,
                                              double* __restrict__ phi_out, double* __restrict__ c_out,
                                              double* __restrict__ mu_out, int rows, double c_gamma, DevScal* S) {
    const int n = P.n;
    bool bad_mix = false, bad_c = false;
    PEARL_CELL_LOOP(KK)
    const int y = i / P.nx;
    double p[3], c;
    if (y + rows < P.ny) {
        const int j = i + rows * P.nx;
        load_phi(phi_in, n, j, p);
        c = c_in[j];
    } else {
        p[0] = 0;
        p[1] = 0;
        p[2] = 1;
        c = c_gamma;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) phi_out[a * n + i] = p[a];
    c_out[i] = c;
    Mix m;
    if (!mixture(P, p, m)) bad_mix = true;
    if (!__builtin_isfinite(c)) bad_c = true;
    mu_out[i] = (c - m.c0) / m.s;
    PEARL_CELL_LOOP_END
    if (bad_mix) atomicOr(&S->err, ERRF_MIXTURE);
    if (bad_c) atomicOr(&S->err, ERRF_NONFINITE_CARBON);
}

// Clears the error bits (a kernel rather than a memset so that the call is
// safe inside captured graphs).
__global__ void k_clear_flags(DevScal* S) { S->err = 0; }

// Marks an attempt live without a phase proposal (implicit_diffusion on a
// caller-supplied trial phi).
__global__ void k_begin_solve(DevScal* S) {
    S->live = 1;
    S->cg_state = CG_IDLE;
}

}  // namespace pearl_gpu
