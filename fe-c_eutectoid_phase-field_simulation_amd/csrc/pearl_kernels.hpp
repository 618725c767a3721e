// Device side of the per-step solver: per-domain parameter blocks, the
// device-resident step controller, pointwise physics, the blocked
// reduction, and the kernels.
//
// Each per-cell expression keeps the reference's operation order and
// rounding (pearl_serial.cpp lines cited per function); the library is built
// with -ffp-contract=off so no multiply-add is fused unless the reference
// calls std::fma.  Loop-invariant quotients that the reference recomputes
// per call (1/(2A), -0.25/A, 16 sigma/(pi^2 eps), ...) are evaluated once on
// the host with the same expression, which yields the same bits.
//
// Control.  advance() (:1119-1174) runs on the device as a small state
// machine per domain (DevScal::phase).  The host enqueues fixed "rounds" of
// kernels -- state reductions, propose, prepare (setup + initial residual),
// PCG slots, finalize, then the decision and the start of the next attempt
// (the turn, taken by the round's last finisher; a batch's first round also
// starts with k_begin) -- and every
// kernel acts only on the domains whose phase and PCG state call for it.  A round therefore has no host decision in it:
// it can be replayed from a captured graph, and one launch serves any number
// of independent domains of different sizes.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace pearl_gpu {

constexpr int TPB = 256;        // threads per block of every cell kernel
constexpr int MAX_SLOTS = 64;   // per-launch work records (profiling)
constexpr int NKIND = 16;       // per-domain run counters

// Per-domain parameter block.  Geometry of the domain inside the packed
// arena, then the model constants, all derived on the host exactly as the
// reference forms them.
struct KP {
    int nx, ny, n;          // grid; n = nx*ny cells
    int nb;                 // blocks (tiles) of this domain in a cell-kernel grid
    int tiles_x;            // tiles per row of tiles: ceil(nx/256)
    int boff;               // first block of this domain in the grid
    long long off;          // first cell of this domain in every plane
    long long poff;         // start of this domain's reduction partials
    double dx, dy, dx2, dy2;  // dx2 = dx*dx (the reference's d.dx*d.dx)
    // 1/d rounded to nearest for every constant divisor d (div_const)
    double rdx, rdy, rdx2, rdy2, rVm, rtwoA[3], r3;
    double A[3], B[3], C[3], D[3];
    double chi[3];          // 1.0/(2.0*A)        Parabola::chi   :84
    double qpsi[3];         // -0.25/A             Parabola::psi   :87
    double twoA[3];         // 2.0*A               Parabola::c_at_mu :83
    double obs, tr;         // 16*sigma/(pi*pi*eps), triple/eps   :964
    double face_coef;       // 2*eps*sigma                        :981
    double eps_sigma;       // eps*sigma                          :816
    // per phase pair (a<b: 01, 02, 12): 16 sigma_ab/(pi^2 eps), 2 eps sigma_ab,
    // eps sigma_ab; all equal to obs / face_coef / eps_sigma unless `pair`
    int pair;               // pairwise interface energies in use
    double sig_p[3], obs_p[3], fc_p[3], es_p[3];
    double eps, sigma, Vm;
    double tau[3], tau_min;
    // FP32 copies for the FP32-stencil phase proposal (propose_z32): the
    // constants above rounded once, and the reciprocals of Vm, dx^2, dy^2
    float f_obs, f_tr, f_obs_p[3], f_fc_p[3], f_rVm, f_rdx2, f_rdy2, f_tau[3], f_tau_min, f_eps;
    double psiref[3];       // psi_a(mu_ref) - psi_gamma(mu_ref), psiref[2]=0  :797
    double mu_ref;
    // Controls (:767-773)
    double dt_max, dt_min, max_phase_change, cg_rtol, cg_atol, closure_tol, energy_rtol, energy_atol;
    double mass_floor;      // gx*gy*dx*dy*1e-6                         :1152
    double energy_scale;    // sigma*max(gx*dx, gy*dy)                  :1158
    int cg_max, max_retries, check_energy, phase_only;
};

// Controller phases of one domain.  PH_PREPARE: the trial phases are given
// (the single implicit_diffusion entry) and the solve starts at prepare.
enum { PH_IDLE = 0, PH_START = 1, PH_PROPOSE = 2, PH_SOLVE = 3, PH_DECIDE = 4, PH_PREPARE = 5 };
// PCG states (implicit_diffusion :1047-1080).  CG_ITER: take a step;
// CG_TRUE: recompute the true residual (:1053) in a pass of its own;
// CG_CHECK: the recursive residual passed, and the finalize pass computes
// the true residual beside the carbon update it would keep (the update is
// discarded when the check fails); CG_OK / CG_FAIL: done; CG_IDLE: not
// started.
enum { CG_IDLE = 0, CG_ITER = 1, CG_TRUE = 2, CG_OK = 3, CG_FAIL = 4, CG_CHECK = 5 };
// Device error codes (the C ABI's PEARL_ERR_* values).
enum { ST_OK = 0, ST_INVALID_TIME = 1, ST_DT_MIN = 2, ST_RETRY_LIMIT = 3, ST_MIXTURE = 4, ST_PROJECTION = 5,
       ST_FIXED_SOLVE = 6 };
enum { ERRF_MIXTURE = 1, ERRF_PROJECTION = 2, ERRF_NONFINITE_CARBON = 4, ERRF_MIXTURE_STATE = 8 };
// Kernels counted per domain (work actually done).
// RUN_CG_APPLY: PCG direction launches that also applied the previous step
// to the solution; RUN_CG_FRESH: those that did not (the first of a solve);
// RUN_FIN_LAG: finalize launches that applied the last step themselves.
enum { RUN_PROPOSE = 0, RUN_PREPARE, RUN_CG_APPLY, RUN_CG_UPDATE, RUN_TRUE_RESIDUAL, RUN_FINALIZE, RUN_STATE,
       RUN_CG_FRESH, RUN_FIN_LAG, RUN_COUNT };

// Device-resident state of one domain: controller, step scalars, the
// results of the reductions and the StepInfo of the last accepted step.
struct DevScal {
    // controller
    int phase, status, why, attempt;
    int cur;               // which state buffer is current
    int need_state;        // e0/mass0 must be computed first
    int e0_valid, m0_valid;
    long long step, rejected, target_steps, taken;
    double time, dt_state, stop_time;  // State::time, State::dt; stop_time of advance
    double stop_eps;       // a run ends once time >= stop_time - stop_eps
    double dt;             // dt of the current attempt
    double e0, m0;
    // attempt
    int live;              // the phase increment passed (:1131)
    int cg_state, cg_it, restart, pidx;
    int cg_alone;          // a solve with no finalize pass after it (the adjoint's): the true residual is a pass of its own
    // propose_phi's reduction: the bits of the largest phase change (a
    // non-negative double orders like its bits), folded by atomicMax into
    // one of 8 slots per block (spread over the XCDs), and the projection
    // failure flag; cleared at the start of every attempt
    unsigned long long chg[8];
    int proj;
    // The forward solve's x = mu - mu_ref: x_mu, the plane is not written yet
    // and x is mu_old - mu_ref; x_lag, x still lacks alpha p of the last
    // step (p in buffer pidx), which the next kernel that needs x applies.
    int x_mu, x_lag;
    int err;               // ERRF_* bits
    double change;
    double rhs2, rz, rr, pAp, alpha, beta, tol, cg_res;
    double closure, valid, mass, energy;
    double e_valid, e_mass, e_energy;
    double front, gmax, removed;
    double dmin_c, dmax_c, dmin_mu, dmax_mu, dsumerr, dmass, denergy;
    // interface smoothing: max |force|, max phase change, energy, failure flag
    double sm_fmax, sm_change, sm_energy, sm_bad;
    // initialization: sums of the carbon intercept and susceptibility, and
    // an invalid-mixture flag
    double in_c0, in_s, in_bad;
    // StepInfo of the last accepted step
    double i_dt, i_change, i_residual, i_closure, i_energy_change;
    int i_cg, i_retries;
    // counters
    long long attempts, rejected_attempts, cg_all, cg_total;
    int cg_peak, pad;
    // fixed-dt mode (the differentiated map of the adjoint): step k takes
    // dt = fixed_dts[k] and is accepted whenever its solve converged
    int fixed, fixed_saved;  // fixed_saved: steps already copied to fixed_save
    const double* fixed_dts;
    double* fixed_save;      // if set: state after fixed step k -> fixed_save + (k-1)*5n
    unsigned runs[NKIND];
};

__device__ __forceinline__ double dmaxd(double a, double b) { return (a < b) ? b : a; }  // std::max
__device__ __forceinline__ double dmind(double a, double b) { return (b < a) ? b : a; }  // std::min

// a/d rounded to nearest, for a positive divisor d fixed at set-up, given
// y = 1/d rounded to nearest.  q0 = a*y is within 2 ulp of a/d.  One
// correction q1 = q0 + (a - q0 d) y (each step one rounding) lands within
// (1/2 + 2^-51) ulp of a/d, so q1 is a faithful rounding.  For a faithful q
// and a reciprocal with relative error below 2^-53, Markstein's theorem
// (P. Markstein, IBM J. Res. Dev. 34(1), 1990) gives RN(q + (a - d q) y) =
// RN(a/d): the second correction returns the IEEE quotient.  The argument
// needs the remainders a - q d exact (|a| >= 2^-969) and the quotient normal
// and finite.  The host admits only divisors in [2^-150, 2^150], so both
// hold when a = 0 or 2^-818 <= |a| < 2^750: frexp's exponent of a in
// [-817, 750] (0 for a = 0).  Any other numerator -- subnormal, tiny,
// huge -- takes the IEEE division itself.  An infinite or NaN numerator
// (frexp exponent 0) stays on the fast path, and div_fixup turns its result
// into the IEEE one (inf with the sign of a, NaN); for every other a it
// returns |q2| with the sign of a (d > 0), which also gives a zero
// numerator its sign back.  In range: one multiply, four FMAs and the
// fix-up instead of the scaled reciprocal iteration of a general division.
__device__ __forceinline__ double div_markstein(double a, double d, double y) {
    const double q0 = a * y;
    const double q1 = __builtin_fma(__builtin_fma(-q0, d, a), y, q0);
    const double q2 = __builtin_fma(__builtin_fma(-q1, d, a), y, q1);
    return __builtin_amdgcn_div_fixup(q2, d, a);
}

__device__ __forceinline__ double div_const(double a, double d, double y) {
    const int e = __builtin_amdgcn_frexp_exp(a);
    if (__builtin_expect(static_cast<unsigned>(e + 817) > 1567u, 0)) return a / d;
    return div_markstein(a, d, y);
}

// The same quotients for a run of constant divisions whose numerators are
// range-checked once, after the run: DivRun takes div_markstein for every
// quotient and keeps the least and greatest frexp exponent of the
// numerators; where ok() fails the caller computes the run again with
// DivIEEE, whose IEEE quotients equal div_markstein's wherever the range
// holds.  In a run of n divisions this costs one exponent, one min and one
// max per division and a single branch, instead of a branch per division.
struct DivRun {
    int lo = 0, hi = 0;
    __device__ __forceinline__ double operator()(double a, double d, double y) {
        const int e = __builtin_amdgcn_frexp_exp(a);
        lo = e < lo ? e : lo;
        hi = e > hi ? e : hi;
        return div_markstein(a, d, y);
    }
    __device__ __forceinline__ bool ok() const { return lo >= -817 && hi <= 750; }
};
struct DivIEEE {
    __device__ __forceinline__ double operator()(double a, double d, double) const { return a / d; }
};
// div_const itself as a policy (each quotient range-checked on its own).
struct DivConst {
    __device__ __forceinline__ double operator()(double a, double d, double y) const { return div_const(a, d, y); }
};

// f(div) once with DivRun, and again with DivIEEE where the run's range
// check fails; f writes its results, so the second call replaces them.
template <class F>
__device__ __forceinline__ void checked_run(F&& f) {
    DivRun fast;
    f(fast);
    if (__builtin_expect(!fast.ok(), 0)) {
        DivIEEE ieee;
        f(ieee);
    }
}

// ---------------------------------------------------------------------------
// Blocked reduction.  OPS packs 2 bits per quantity: 0 sum, 1 max, 2 min.
// Stage 1 (caller): each thread folds its cells into v[q] in cell order.
// Stage 2: 64-lane xor-halving tree, then (w0+w1)+(w2+w3) per block; block
// partials to part[q*nb+b].  Stage 3, in a separate finisher kernel: the
// partials are reduced in the same shape.  Splitting stages 2 and 3 across
// a kernel boundary avoids device-scope fences in every block (on a
// multi-XCD part such a fence writes back the XCD's L2).
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

// Stage 2 inside a cell kernel: the block's partials (thread 0 writes them).
// Blocks larger than TPB (cg_update's) hold the values in their first TPB
// threads; the other waves only take part in the barrier.
template <int NQ, unsigned OPS>
__device__ __forceinline__ void block_partial(double (&v)[NQ], double* __restrict__ part, int nb, int lb) {
    __shared__ double sh[4][NQ];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (threadIdx.x < TPB) {
        wave_tree<NQ, OPS>(v);
        if (lane == 0) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) sh[w][q] = v[q];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int op = op_of<OPS>(q);
            part[q * nb + lb] = comb(op, comb(op, sh[0][q], sh[1][q]), comb(op, sh[2][q], sh[3][q]));
        }
    }
}

// Stage 3 in a finisher kernel (one 256-thread block per domain, launched
// after the cell kernel, so its partials are visible without fences):
// thread t folds partials t, t+256, ... then the same tree.  Returns true in
// thread 0, with the totals in v.  The partials are fetched FOLD at a time
// ahead of the fold, so a thread has that many loads in flight instead of
// one dependent load per partial.
template <int NQ, unsigned OPS>
__device__ __forceinline__ bool final_reduce(double (&v)[NQ], const double* __restrict__ part, int nb) {
    constexpr int FOLD = NQ <= 4 ? 32 : 16;
    __shared__ double sh[4][NQ];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) v[q] = ident(op_of<OPS>(q));
    for (int b0 = static_cast<int>(threadIdx.x); b0 < nb; b0 += FOLD * TPB) {
        double t[NQ][FOLD];
#pragma unroll
        for (int u = 0; u < FOLD; ++u) {
            const int b = b0 + u * TPB;
#pragma unroll
            for (int q = 0; q < NQ; ++q) t[q][u] = b < nb ? part[q * nb + b] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < FOLD; ++u) {
            if (b0 + u * TPB >= nb) break;
#pragma unroll
            for (int q = 0; q < NQ; ++q) v[q] = comb(op_of<OPS>(q), v[q], t[q][u]);
        }
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
    return true;
}

// ---------------------------------------------------------------------------
// Pointwise physics (pearl_serial.cpp:76-145, :533-537, :491-509)
// ---------------------------------------------------------------------------
// Index of the phase pair a < b: 01 -> 0, 02 -> 1, 12 -> 2.
__device__ __forceinline__ constexpr int pair_of(int a, int b) { return a + b - 1; }

// Obstacle force on phase a: the reference's obs (phi_a1 + phi_a2) (:964),
// or with pairwise interface energies obs_{a,a1} phi_a1 + obs_{a,a2} phi_a2.
__device__ __forceinline__ double obstacle_force(const KP& P, const double (&p)[3], int a) {
    const int a1 = (a + 1) % 3, a2 = (a + 2) % 3;
    if (P.pair) return P.obs_p[pair_of(a < a1 ? a : a1, a < a1 ? a1 : a)] * p[a1] +
                       P.obs_p[pair_of(a < a2 ? a : a2, a < a2 ? a2 : a)] * p[a2];
    return P.obs * (p[a1] + p[a2]);
}

// Obstacle energy density: obs sum_{a<b} phi_a phi_b (:802), or with
// pairwise interface energies sum_{a<b} obs_ab phi_a phi_b.
__device__ __forceinline__ double obstacle_energy(const KP& P, const double (&p)[3]) {
    if (P.pair) return P.obs_p[0] * (p[0] * p[1]) + P.obs_p[1] * (p[0] * p[2]) + P.obs_p[2] * (p[1] * p[2]);
    return P.obs * (p[0] * p[1] + p[0] * p[2] + p[1] * p[2]);
}

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

// project_simplex :491-509; false where the reference throws.  FINITE: the
// caller has checked that z is finite, and the descending sort takes max and
// min instead of the reference's compare-and-swap.  For finite values the
// two can differ only in the zeros among the sorted values (a +0 against a
// -0), and those change neither a partial sum's value before 1 is
// subtracted nor any comparison, so the threshold and the result are the
// same.
template <bool FINITE = false, class Div>
__device__ __forceinline__ bool project_simplex(const double (&z)[3], double r3, double (&out)[3], Div& div) {
    double s0 = z[0], s1 = z[1], s2 = z[2], t;
    if constexpr (FINITE) {
        t = s0;
        s0 = __builtin_fmax(t, s1);
        s1 = __builtin_fmin(t, s1);
        t = s1;
        s1 = __builtin_fmax(t, s2);
        s2 = __builtin_fmin(t, s2);
        t = s0;
        s0 = __builtin_fmax(t, s1);
        s1 = __builtin_fmin(t, s1);
    } else {
        if (s0 < s1) { t = s0; s0 = s1; s1 = t; }
        if (s1 < s2) { t = s1; s1 = s2; s2 = t; }
        if (s0 < s1) { t = s0; s0 = s1; s1 = t; }
    }
    double sum = 0.0, threshold = 0.0;
    int active = 0;
    sum += s0;
    { const double cand = (sum - 1.0) / 1.0; if (s0 > cand) { active = 1; threshold = cand; } }
    sum += s1;
    { const double cand = (sum - 1.0) / 2.0; if (s1 > cand) { active = 2; threshold = cand; } }
    sum += s2;
    { const double cand = div(sum - 1.0, 3.0, r3); if (s2 > cand) { active = 3; threshold = cand; } }
#pragma unroll
    for (int a = 0; a < 3; ++a) out[a] = dmaxd(0.0, z[a] - threshold);
    return active != 0;
}

__device__ __forceinline__ bool project_simplex(const double (&z)[3], double r3, double (&out)[3]) {
    DivConst div;
    return project_simplex(z, r3, out, div);
}

// Stencil neighbours of cell i in row y: x periodic; at the y edges the
// neighbour index falls back to the cell itself (the reference's y ghosts
// are copies of the edge row, :692-693).
struct Nbr {
    int y, w, e, s, nn;
};

__device__ __forceinline__ void load_phi(const double* __restrict__ phi, int n, int i, double (&p)[3]) {
    p[0] = phi[i];
    p[1] = phi[n + i];
    p[2] = phi[2 * n + i];
}

// weights_at :1022-1028, from the diffusivity of the cell and its four
// neighbours (w, e, s, n).
__device__ __forceinline__ void face_weights(const KP& P, int y, double k, double kw, double ke, double ks, double kn,
                                             double (&w)[4]) {
    checked_run([&](auto& div) {
        w[0] = div(.5 * (k + kw), P.dx2, P.rdx2);
        w[1] = div(.5 * (k + ke), P.dx2, P.rdx2);
        w[2] = div(.5 * (k + ks), P.dy2, P.rdy2);
        w[3] = div(.5 * (k + kn), P.dy2, P.rdy2);
    });
    if (y == 0) w[2] = 0;
    if (y == P.ny - 1) w[3] = 0;
}

// apply :1029-1035 for one cell: chi v + dt sum_f w_f (v - v_f).
__device__ __forceinline__ double apply_vals(double chi, double dt, const double (&w)[4], double v, double vw, double ve,
                                             double vs, double vn) {
    return chi * v + dt * (w[0] * (v - vw) + w[1] * (v - ve) + w[2] * (v - vs) + w[3] * (v - vn));
}

// K of mixture (:137) alone, in the same operation order.
__device__ __forceinline__ double mixture_K(const KP& P, const double (&phi)[3]) {
    double K = 0.0;
#pragma unroll
    for (int p = 0; p < 3; ++p) K += phi[p] * P.D[p] * P.chi[p];
    return K;
}

// Per-cell acceptance quantities of valid_state (:1098-1114) and the local
// energy of reduced_energy (:799-819): the cell term, then the east and
// (below the top row) north face terms in the reference's order.  p, qe, qn:
// phases of the cell and of its east and north neighbours.  `known`: the
// cell's mixture when the caller has already formed it (the same value).
// div: the constant-division policy; accept_vals below runs it checked.
template <class Div>
__device__ __forceinline__ void accept_vals_d(const KP& P, const double (&p)[3], const double (&qe)[3],
                                              const double (&qn)[3], bool north, double c, double mu, bool& good,
                                              double& energy, bool& mix_ok, const Mix* known, Div& div) {
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
                const double cp = div(mu - P.B[a], P.twoA[a], P.rtwoA[a]);
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
        if (known) m = *known;
        else if (!mixture(P, p, m)) mix_ok = false;
        const double dc = c - composition(P.mu_ref, m);
        chem += .5 * dc * dc / m.s;
    }
    double v = div(chem, P.Vm, P.rVm) + obstacle_energy(P, p) + P.tr * p[0] * p[1] * p[2];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int bb = a + 1; bb < 3; ++bb) {
            const double z = div(p[a] * qe[bb] - qe[a] * p[bb], P.dx, P.rdx);
            v += P.es_p[pair_of(a, bb)] * z * z;
        }
    if (north) {
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int bb = a + 1; bb < 3; ++bb) {
                const double z = div(p[a] * qn[bb] - qn[a] * p[bb], P.dy, P.rdy);
                v += P.es_p[pair_of(a, bb)] * z * z;
            }
    }
    energy = v;
}

// accept_vals_d with the cell's constant divisions as one checked run; good
// and mix_ok are cleared where the cell fails.
__device__ __forceinline__ void accept_vals(const KP& P, const double (&p)[3], const double (&qe)[3],
                                            const double (&qn)[3], bool north, double c, double mu, bool& good,
                                            double& energy, bool& mix_ok, const Mix* known = nullptr) {
    bool g, mo;
    checked_run([&](auto& div) {
        g = true;
        mo = true;
        accept_vals_d(P, p, qe, qn, north, c, mu, g, energy, mo, known, div);
    });
    good = good && g;
    mix_ok = mix_ok && mo;
}

__device__ __forceinline__ void accept_cell(const KP& P, const double* __restrict__ phi, int i, const Nbr& b,
                                            double c, double mu, bool& good, double& energy, bool& mix_ok) {
    double p[3], qe[3], qn[3];
    load_phi(phi, P.n, i, p);
    load_phi(phi, P.n, b.e, qe);
    const bool north = b.y != P.ny - 1;
    if (north) load_phi(phi, P.n, b.nn, qn);
    else load_phi(phi, P.n, i, qn);
    accept_vals(P, p, qe, qn, north, c, mu, good, energy, mix_ok);
}

// ---------------------------------------------------------------------------
// Packed arena and domain lookup
// ---------------------------------------------------------------------------
struct Bufs {
    double* phi[2];   // domain d: phi[b] + 3*off + a*n + i
    double* c[2];
    double* mu[2];
    double *chi, *K, *rhs, *x, *diag, *r, *z, *p0, *p1, *Ap;
    double* part;     // reduction partials, domain d at part + poff
    const KP* kp;
    DevScal* S;
    const int* blk2dom;
    int* kinds;       // per-launch work records (profiling; may be unused)
    int ndom;
};

struct Where {
    int d, lb;
};

// Domain and block-in-domain of this block: a round grid covers all
// domains; a direct launch (only_d >= 0) covers one domain.
// The domain index is made wave-uniform (readfirstlane), so the domain's
// parameters are read through the scalar cache; kernels copy their KP by
// value, which keeps the fields in registers instead of re-reading them
// after every global store (the stores could alias the parameter block).
// A single-domain context skips the table: its one domain starts at block 0,
// so the block's controller state is the first load it waits for (a
// predicated-off launch exits one load latency sooner).
__device__ __forceinline__ Where where(const Bufs& B, int only_d) {
    const int g = static_cast<int>(blockIdx.x);
    if (only_d >= 0) return {only_d, g};
    if (B.ndom == 1) return {0, g};
    const int d = __builtin_amdgcn_readfirstlane(B.blk2dom[g]);
    return {d, g - B.kp[d].boff};
}

__device__ __forceinline__ void record(const Bufs& B, int slot, int lb, int kind) {
    if (slot >= 0 && lb == 0 && threadIdx.x == 0) B.kinds[slot] = kind;
}

// ---------------------------------------------------------------------------
// Tiles.  Block lb of a domain covers the tile of 256 columns by TY rows at
// tile column lb % tiles_x, tile row lb / tiles_x.  Thread t owns column
// 256*(lb % tiles_x) + t and visits its cells of the tile in increasing row:
// that is the order of each thread's fold in the blocked reduction.  Down a
// column consecutive cells are one row apart, so a stencil kernel keeps the
// south, centre and north values of its column in registers and loads one
// new row per cell; the west and east neighbours are loads from the lines
// that the neighbouring lanes fetch for their own centres.
// ---------------------------------------------------------------------------
struct Col {
    int x, y0, y1;  // column; rows [y0, y1)
    int i0;         // first cell
    int dw, de;     // offsets of the west and east neighbours (x periodic)
    bool on;        // the column lies inside the domain
};

template <int TY>
__device__ __forceinline__ Col column(const KP& P, int lb, int rows, int t = -1) {
    Col c;
    const int tx = lb % P.tiles_x, ty = lb / P.tiles_x;
    c.x = tx * TPB + (t < 0 ? static_cast<int>(threadIdx.x) : t);
    c.y0 = ty * TY;
    c.y1 = c.y0 + TY < rows ? c.y0 + TY : rows;
    c.on = c.x < P.nx;
    c.i0 = c.y0 * P.nx + c.x;
    c.dw = c.x == 0 ? P.nx - 1 : -1;
    c.de = c.x == P.nx - 1 ? 1 - P.nx : 1;
    return c;
}

__device__ __forceinline__ Nbr nbr_of(const KP& P, const Col& c, int i, int y) {
    Nbr r;
    r.y = y;
    r.w = i + c.dw;
    r.e = i + c.de;
    r.s = y == 0 ? i : i - P.nx;
    r.nn = y == P.ny - 1 ? i : i + P.nx;
    return r;
}

// Visit the column's cells in order: body(i, y).
template <int TY, class F>
__device__ __forceinline__ void for_column(const KP& P, const Col& c, F&& body) {
    if (!c.on) return;
    int i = c.i0;
#pragma unroll
    for (int r = 0; r < TY; ++r) {
        const int y = c.y0 + r;
        if (y >= c.y1) break;
        body(i, y);
        i += P.nx;
    }
}

// March down the column with the south, centre and north values (type T,
// produced by load(cell)) in registers: body(i, y, south, centre, north).
// At the bottom and top rows the missing neighbour is the cell itself.
template <int TY, class T, class L, class F>
__device__ __forceinline__ void march(const KP& P, const Col& c, L&& load, F&& body) {
    if (!c.on) return;
    int i = c.i0;
    T s = load(c.y0 == 0 ? i : i - P.nx);
    T m = load(i);
#pragma unroll
    for (int r = 0; r < TY; ++r) {
        const int y = c.y0 + r;
        if (y >= c.y1) break;
        const T n = load(y == P.ny - 1 ? i : i + P.nx);
        body(i, y, s, m, n);
        s = m;
        m = n;
        i += P.nx;
    }
}

// Streaming accesses: a value read once (no stencil reuse) or written for a
// later kernel moves with the nontemporal hint, so it does not displace the
// stencil rows that neighbouring columns and tiles re-read from L2.
template <class T>
__device__ __forceinline__ T ld_once(const T* p) {
    return __builtin_nontemporal_load(p);
}
template <class T>
__device__ __forceinline__ void st_once(T* p, T v) {
    __builtin_nontemporal_store(v, p);
}

// Loop head of implicit_diffusion (:1047-1067) for iteration `it` with the
// current r.r: info.it, info.residual, then the convergence and
// iteration-limit branches.  One thread.
__device__ __forceinline__ void cg_head(const KP& P, DevScal* S, int it) {
    const double res = __builtin_sqrt(S->rr);
    S->cg_it = it;
    S->cg_res = res;
    if (!__builtin_isfinite(res)) S->cg_state = CG_FAIL;
    else if (res <= S->tol) S->cg_state = S->cg_alone ? CG_TRUE : CG_CHECK;
    else if (it == P.cg_max) S->cg_state = CG_FAIL;
    else S->cg_state = CG_ITER;
    if (S->cg_state == CG_FAIL) S->phase = PH_DECIDE;
}

// ---------------------------------------------------------------------------
// Round kernels
// ---------------------------------------------------------------------------

// Start of a step (:1120-1123) and reset of the per-attempt flags, for
// domain d.
__device__ __forceinline__ void begin_domain(const Bufs& B, int d) {
    const KP& P = B.kp[d];
    DevScal* S = B.S + d;
    if (S->phase == PH_START) {
        double dt;
        if (S->fixed) {
            dt = S->fixed_dts[S->taken];
        } else {
            dt = S->dt_state;
            if (P.dt_max < dt) dt = P.dt_max;
            if (S->stop_time - S->time < dt) dt = S->stop_time - S->time;
        }
        if (!(dt > 0) || !__builtin_isfinite(dt)) {
            S->status = ST_INVALID_TIME;
            S->phase = PH_IDLE;
            return;
        }
        S->dt = dt;
        S->attempt = 0;
        S->phase = PH_PROPOSE;
        S->need_state =
            (!S->fixed && ((P.check_energy && !S->e0_valid) || (!P.phase_only && !S->m0_valid))) ? 1 : 0;
    }
    if (S->phase == PH_PROPOSE) {
#pragma unroll
        for (int k = 0; k < 8; ++k) S->chg[k] = 0;
        S->proj = 0;
        S->live = 0;
        S->cg_alone = 0;
        S->cg_state = CG_IDLE;
        S->cg_it = 0;
        S->cg_res = 0;
        S->closure = 0;
        S->restart = 1;
        S->pidx = 0;
        S->pad = 0;  // "stepped": k_cg_a took a step that k_cg_b must complete
    }
}

// One thread per domain.
__global__ void k_begin(Bufs B) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d < B.ndom) begin_domain(B, d);
}

// valid_state / total_carbon / reduced_energy of one cell of a state
// (:1096-1116, :784-788, :793-821).
__device__ __forceinline__ void state_cell(const KP& P, const double* phi, const double* c, const double* mu, int i,
                                           const Nbr& nb, bool& good, double& mass, double& energy, bool& mix_ok) {
    double e;
    const double ci = c[i];
    accept_cell(P, phi, i, nb, ci, mu[i], good, e, mix_ok);
    mass += ci;
    energy += e;
}

// e0 and mass0 of a step (:1122-1123) when they are not cached.
template <int TY>
__global__ __launch_bounds__(TPB) void k_state(Bufs B, int slot) {
    const Where w = where(B, -1);
    const KP P = B.kp[w.d];
    DevScal* S = B.S + w.d;
    if (!(S->phase == PH_PROPOSE && S->need_state)) return;
    record(B, slot, w.lb, 1);
    const int cur = S->cur;
    const double* phi = B.phi[cur] + 3 * P.off;
    const double* c = B.c[cur] + P.off;
    const double* mu = B.mu[cur] + P.off;
    double v[4] = {1.0, 0.0, 0.0, 0.0};
    bool good = true, mix_ok = true;
    const Col col = column<TY>(P, w.lb, P.ny);
    for_column<TY>(P, col, [&](const int i, const int y) {
        state_cell(P, phi, c, mu, i, nbr_of(P, col, i, y), good, v[1], v[2], mix_ok);
    });
    v[0] = good ? 1.0 : 0.0;
    v[3] = mix_ok ? 0.0 : 1.0;
    block_partial<4, (2u | (1u << 6))>(v, B.part + P.poff, P.nb, w.lb);
}

__global__ __launch_bounds__(TPB) void k_state_fin(Bufs B) {
    const int d = blockIdx.x;
    const KP P = B.kp[d];
    DevScal* S = B.S + d;
    if (!(S->phase == PH_PROPOSE && S->need_state)) return;
    double v[4];
    if (!final_reduce<4, (2u | (1u << 6))>(v, B.part + P.poff, P.nb)) return;
    S->runs[RUN_STATE]++;
    if (P.check_energy && v[3] > 0) {
        S->status = ST_MIXTURE;
        S->phase = PH_IDLE;
        return;
    }
    if (P.check_energy) {
        S->e0 = v[2] * P.dx * P.dy;
        S->e0_valid = 1;
    }
    if (!P.phase_only) {
        S->m0 = v[1] * P.dx * P.dy;
        S->m0_valid = 1;
    }
    S->need_state = 0;
}

// Trial value of propose_phi for one cell (:963-996): the force, its mean,
// the step scale and z = phi - scale*(force - mean), in the reference's
// operation order.  p and its west, east, south, north neighbours qw, qe,
// qs, qn (the south/north terms are skipped at the y edges).
// The face term w of a pair on a cell's south face is the negation of the
// term on its southern neighbour's north face (the difference, the product
// and the constant division are all sign-symmetric in IEEE arithmetic),
// except that where the two products are equal both differences are +0, so
// the negation is -0 where the direct term is +0.  That sign never shows:
// adding or subtracting a signed zero leaves the force unchanged unless the
// force is -0, and it never is.  It starts as obs (phi_a1 + phi_a2) (or
// obs_a,a1 phi_a1 + obs_a,a2 phi_a2) plus tr phi_a1 phi_a2 with obs > 0 and
// tr >= 0; the first part is -0 only when both phases are -0 or negative,
// which makes the second +0 or positive; and a sum or difference is -0 only
// when the force already is.  So a column march passes the north terms
// down: `south`, when given, holds the southern neighbour's north terms,
// and `north` receives this cell's.  div: the constant-division policy
// (DivRun or DivIEEE).
template <class Div>
__device__ __forceinline__ void propose_z(const KP& P, int y, const double (&p)[3], const double (&qw)[3],
                                          const double (&qe)[3], const double (&qs)[3], const double (&qn)[3],
                                          double m, double dt, double (&force)[3], double& mean, double& scale,
                                          double (&z)[3], const double* south, double (&north)[3], Div& div) {
    const double psi0 = psi(P, 2, m);
    const double dtr[3] = {2.0 * p[1] * p[2], 2.0 * p[0] * p[2], 2.0 * p[0] * p[1]};
    double dpsi[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) dpsi[b] = psi(P, b, m) - psi0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int a1 = (a + 1) % 3, a2 = (a + 2) % 3;
        force[a] = obstacle_force(P, p, a) + P.tr * p[a1] * p[a2];
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const double dh = dtr[a] + (b == a ? 6.0 * p[b] * (1.0 - p[b]) : 0.0);
            // dpsi[2] = psi0 - psi0 is +0 (NaN if psi0 is not finite), so its
            // product is a signed zero or NaN, which divided by the positive
            // Vm is itself: that quotient needs no division.
            const double t = dpsi[b] * dh;
            force[a] += b == 2 ? t : div(t, P.Vm, P.rVm);
        }
    }
#pragma unroll
    for (int dir = 0; dir < 4; ++dir) {
        if ((dir == 2 && y == 0) || (dir == 3 && y == P.ny - 1)) continue;
        const double(&q)[3] = dir == 0 ? qw : (dir == 1 ? qe : (dir == 2 ? qs : qn));
        const double dd = dir < 2 ? P.dx2 : P.dy2, rdd = dir < 2 ? P.rdx2 : P.rdy2;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = a + 1; b < 3; ++b) {
                const double wv = (dir == 2 && south)
                                      ? -south[pair_of(a, b)]
                                      : div(P.fc_p[pair_of(a, b)] * (p[a] * q[b] - q[a] * p[b]), dd, rdd);
                if (dir == 3) north[pair_of(a, b)] = wv;
                force[a] += wv * q[b];
                force[b] -= wv * q[a];
            }
    }
    mean = div(force[0] + force[1] + force[2], 3.0, P.r3);
    scale = dt / (local_tau(P, p) * P.eps);
#pragma unroll
    for (int a = 0; a < 3; ++a) z[a] = p[a] - scale * (force[a] - mean);
}

// The FP32-stencil form of propose_z (precision PEARL_FP32_STENCIL): the
// same terms in the same order, with the force, its mean and the step scale
// in FP32.  The chemical driving-force differences psi_b - psi_gamma are
// formed in FP64 (each psi is ~1e4..1e5 J/mol, their differences ~1e1..1e2,
// which FP32 would cancel to 1e-4 relative) and then rounded; the increment
// scale*(force - mean) is rounded once and subtracted from the FP64 phases,
// so the projection, the phase-increment test and the stored phases stay
// FP64.  Constant divisions are multiplications by FP32 reciprocals.
__device__ __forceinline__ void propose_z32(const KP& P, int y, const double (&pd)[3], const double (&qwd)[3],
                                            const double (&qed)[3], const double (&qsd)[3], const double (&qnd)[3],
                                            double m, double dt, double (&z)[3], const float* south,
                                            float (&north)[3]) {
    const double psi0 = psi(P, 2, m);
    float dpsi[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) dpsi[b] = static_cast<float>(psi(P, b, m) - psi0);
    float p[3], qw[3], qe[3], qs[3], qn[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        p[a] = static_cast<float>(pd[a]);
        qw[a] = static_cast<float>(qwd[a]);
        qe[a] = static_cast<float>(qed[a]);
        qs[a] = static_cast<float>(qsd[a]);
        qn[a] = static_cast<float>(qnd[a]);
    }
    const float dtr[3] = {2.0f * p[1] * p[2], 2.0f * p[0] * p[2], 2.0f * p[0] * p[1]};
    float force[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int a1 = (a + 1) % 3, a2 = (a + 2) % 3;
        const float obs = P.pair ? P.f_obs_p[pair_of(a < a1 ? a : a1, a < a1 ? a1 : a)] * p[a1] +
                                       P.f_obs_p[pair_of(a < a2 ? a : a2, a < a2 ? a2 : a)] * p[a2]
                                 : P.f_obs * (p[a1] + p[a2]);
        force[a] = obs + P.f_tr * p[a1] * p[a2];
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const float dh = dtr[a] + (b == a ? 6.0f * p[b] * (1.0f - p[b]) : 0.0f);
            force[a] += dpsi[b] * dh * P.f_rVm;
        }
    }
#pragma unroll
    for (int dir = 0; dir < 4; ++dir) {
        if ((dir == 2 && y == 0) || (dir == 3 && y == P.ny - 1)) continue;
        const float(&q)[3] = dir == 0 ? qw : (dir == 1 ? qe : (dir == 2 ? qs : qn));
        const float rdd = dir < 2 ? P.f_rdx2 : P.f_rdy2;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = a + 1; b < 3; ++b) {
                const float wv = (dir == 2 && south) ? -south[pair_of(a, b)]
                                                     : P.f_fc_p[pair_of(a, b)] * (p[a] * q[b] - q[a] * p[b]) * rdd;
                if (dir == 3) north[pair_of(a, b)] = wv;
                force[a] += wv * q[b];
                force[b] -= wv * q[a];
            }
    }
    const float mean = (force[0] + force[1] + force[2]) * (1.0f / 3.0f);
    const float w0 = p[0] * p[1], w1 = p[0] * p[2], w2 = p[1] * p[2];
    const float ws = w0 + w1 + w2;
    const float tau = ws > 1e-30f ? (w0 * P.f_tau[0] + w1 * P.f_tau[1] + w2 * P.f_tau[2]) / ws : P.f_tau_min;
    const float scale = static_cast<float>(dt) / (tau * P.f_eps);
#pragma unroll
    for (int a = 0; a < 3; ++a) z[a] = pd[a] - static_cast<double>(scale * (force[a] - mean));
}

// propose_phi :961-1002 for one cell: p and the south/north phases from the
// column registers, west/east loaded.  v[0] = max change, v[1] = projection
// failure.  The FP32-stencil form (propose_z32); the carried north face
// terms are FP32.
__device__ __forceinline__ void propose_cell32(const KP& P, const double* __restrict__ phi, int n, int i, int y,
                                               const Col& col, const double (&ps)[3], const double (&p)[3],
                                               const double (&pn)[3], double m, double dt, double* __restrict__ out,
                                               double (&v)[2], const float* south, float (&north)[3]) {
    double qw[3], qe[3];
    load_phi(phi, n, i + col.dw, qw);
    load_phi(phi, n, i + col.de, qe);
    double z[3], zp[3];
    propose_z32(P, y, p, qw, qe, ps, pn, m, dt, z, south, north);
    if (!(__builtin_isfinite(z[0]) && __builtin_isfinite(z[1]) && __builtin_isfinite(z[2]))) {
        v[0] = __builtin_inf();
#pragma unroll
        for (int a = 0; a < 3; ++a) st_once(out + a * n + i, p[a]);
        return;
    }
    if (!project_simplex(z, P.r3, zp)) v[1] = 1.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) v[0] = dmaxd(v[0], __builtin_fabs(zp[a] - p[a]));
#pragma unroll
    for (int a = 0; a < 3; ++a) st_once(out + a * n + i, zp[a]);
}

// The FP64 form of the same, with the constant-division policy `div`.  With
// DivRun it returns whether the cell's range check held; the cell is stored
// and accumulated only when it held and `commit` is set (commit: the south
// terms came from a cell whose check held).  Otherwise the kernel computes
// the cell again with DivIEEE after its march.
template <class Div>
__device__ __forceinline__ bool propose_cell64(const KP& P, const double* __restrict__ phi, int n, int i, int y,
                                               const Col& col, const double (&ps)[3], const double (&p)[3],
                                               const double (&pn)[3], double m, double dt, double* __restrict__ out,
                                               double (&v)[2], const double* south, double (&north)[3], Div& div,
                                               bool commit = true) {
    double qw[3], qe[3];
    load_phi(phi, n, i + col.dw, qw);
    load_phi(phi, n, i + col.de, qe);
    double z[3], zp[3], force[3], mean, scale;
    propose_z(P, y, p, qw, qe, ps, pn, m, dt, force, mean, scale, z, south, north, div);
    const bool finite = __builtin_isfinite(z[0]) && __builtin_isfinite(z[1]) && __builtin_isfinite(z[2]);
    const bool projected = !finite || project_simplex<true>(z, P.r3, zp, div);
    if constexpr (std::is_same<Div, DivRun>::value) {
        const bool held = div.ok();
        if (__builtin_expect(!(held && commit), 0)) return held;
    }
    if (!finite) {
        v[0] = __builtin_inf();
#pragma unroll
        for (int a = 0; a < 3; ++a) st_once(out + a * n + i, p[a]);
        return true;
    }
    if (!projected) v[1] = 1.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) v[0] = dmaxd(v[0], __builtin_fabs(zp[a] - p[a]));
#pragma unroll
    for (int a = 0; a < 3; ++a) st_once(out + a * n + i, zp[a]);
    return true;
}

// LDS tile of a cell kernel: TY rows of the block's 256 columns, the row
// above and the row below (clamped at the domain's bottom and top rows, where
// a missing neighbour is the cell itself), and one column either side
// (periodic in x).  stage_tile calls put(r, cc, j, own) for LDS row r, LDS
// column cc and cell j; each thread stages its own column (own: a
// compile-time true; its rows 1..TY are the thread's cells), threads 0 and
// 255 the two halo columns.  The tile is formed once per block, so a
// neighbour's value is neither loaded again nor recomputed by the cells
// around it.
template <int TY>
struct TileLds {
    double a[TY + 2][TPB + 2], b[TY + 2][TPB + 2];
};

template <int TY, class F>
__device__ __forceinline__ void stage_tile(const KP& P, const Col& col, F&& put) {
    const int t = static_cast<int>(threadIdx.x);
    const int x0 = col.x - t;
    auto one = [&](const int cc, int gx, auto own) {
        gx = gx < 0 ? gx + P.nx : (gx >= P.nx ? gx % P.nx : gx);
#pragma unroll
        for (int r = 0; r < TY + 2; ++r) {
            int y = col.y0 - 1 + r;
            y = y < 0 ? 0 : (y > P.ny - 1 ? P.ny - 1 : y);
            put(r, cc, y * P.nx + gx, own);
        }
    };
    one(t + 1, x0 + t, std::true_type{});
    if (t == 0) one(0, x0 - 1, std::false_type{});
    if (t == TPB - 1) one(TPB + 1, x0 + TPB, std::false_type{});
}

// The block's cells in the march order, after staging: body(i, y, r, cc)
// with LDS row r (south r - 1, north r + 1) and column cc (west cc - 1,
// east cc + 1).
template <int TY, class F>
__device__ __forceinline__ void for_tile(const KP& P, const Col& col, F&& body) {
    if (!col.on) return;
    const int cc = static_cast<int>(threadIdx.x) + 1;
    int i = col.i0;
#pragma unroll
    for (int r = 1; r <= TY; ++r) {
        const int y = col.y0 + r - 1;
        if (y >= col.y1) break;
        body(i, y, r, cc);
        i += P.nx;
    }
}

struct Phases {
    double p[3];
};

// P1: propose_phi :961-1002, then the phase-increment test (:1131).
// F32: the FP32-stencil form.
template <int TY, bool F32>
__global__ __launch_bounds__(TPB) void k_propose(Bufs B, int slot) {
    using W = typename std::conditional<F32, float, double>::type;
    const Where w = where(B, -1);
    const KP P = B.kp[w.d];
    DevScal* S = B.S + w.d;
    if (S->phase != PH_PROPOSE || S->need_state) return;
    record(B, slot, w.lb, 1);
    const int n = P.n;
    const int cur = S->cur;
    const double* __restrict__ phi = B.phi[cur] + 3 * P.off;
    const double* __restrict__ mu = B.mu[cur] + P.off;
    double* __restrict__ out = B.phi[cur ^ 1] + 3 * P.off;
    const double dt = S->dt;
    double v[2] = {0.0, 0.0};
    W carry[3] = {0, 0, 0};
    const Col col = column<TY>(P, w.lb, P.ny);
    // FP64: rows whose constant-division range check failed, or whose south
    // terms came from such a row, redone below
    unsigned redo = 0;
    bool held = true;
    march<TY, Phases>(
        P, col,
        [&](const int j) {
            Phases q;
            load_phi(phi, n, j, q.p);
            return q;
        },
        [&](const int i, const int y, const Phases& s, const Phases& c, const Phases& nn) {
            // the previous cell of the march (y - 1) computed this cell's south terms
            W north[3] = {0, 0, 0};
            if constexpr (F32) {
                propose_cell32(P, phi, n, i, y, col, s.p, c.p, nn.p, ld_once(mu + i), dt, out, v,
                               y > col.y0 ? carry : nullptr, north);
            } else {
                DivRun fast;
                const bool south_held = held;
                held = propose_cell64(P, phi, n, i, y, col, s.p, c.p, nn.p, ld_once(mu + i), dt, out, v,
                                      y > col.y0 ? carry : nullptr, north, fast, south_held);
                if (!(held && south_held)) redo |= 1u << (y - col.y0);
            }
#pragma unroll
            for (int q = 0; q < 3; ++q) carry[q] = north[q];
        });
    if constexpr (!F32) {
        // Cells where a numerator left the fast quotient's range (rare), and
        // the cell above each: every term again with IEEE quotients, the
        // south terms formed directly.  Out of the march, this path adds few
        // registers to it.
        if (__builtin_expect(redo != 0, 0)) {
            for (int r = 0; r < TY; ++r) {
                if (!(redo & (1u << r))) continue;
                const int y = col.y0 + r, i = col.i0 + r * P.nx;
                double ps[3], pc[3], pn[3], north[3];
                load_phi(phi, n, y == 0 ? i : i - P.nx, ps);
                load_phi(phi, n, i, pc);
                load_phi(phi, n, y == P.ny - 1 ? i : i + P.nx, pn);
                DivIEEE ieee;
                propose_cell64(P, phi, n, i, y, col, ps, pc, pn, ld_once(mu + i), dt, out, v, nullptr, north, ieee);
            }
        }
    }
    // the block's largest change and projection flag (the same tree as the
    // blocked reduction; a maximum is exact in any order), then one atomic
    // per block: no partials and no finisher launch
    {
        __shared__ double sh[4][2];
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        wave_tree<2, (1u | (1u << 2))>(v);
        if (lane == 0) {
            sh[wv][0] = v[0];
            sh[wv][1] = v[1];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const double m = dmaxd(dmaxd(sh[0][0], sh[1][0]), dmaxd(sh[2][0], sh[3][0]));
            // a zero maximum changes no slot, so a block of unchanged cells
            // (a colony's bulk) issues no atomic: same-address atomics from
            // every block, arriving together, serialize past the XCDs' L2s
            const unsigned long long bits = static_cast<unsigned long long>(__double_as_longlong(m));
            if (bits != 0) atomicMax(&S->chg[blockIdx.x & 7], bits);
            if (dmaxd(dmaxd(sh[0][1], sh[1][1]), dmaxd(sh[2][1], sh[3][1])) > 0) atomicOr(&S->proj, 1);
        }
    }
}

// propose_phi's result for a domain (after the propose launch): the
// largest phase change and whether the attempt goes on (:1131; a fixed-dt
// step always continues); 2 = a projection failed.
__device__ __forceinline__ int propose_verdict(const KP& P, const DevScal* S, double& change) {
    unsigned long long b = S->chg[0];
#pragma unroll
    for (int k = 1; k < 8; ++k) b = S->chg[k] > b ? S->chg[k] : b;
    change = __longlong_as_double(static_cast<long long>(b));
    if (S->proj) return 2;
    return (S->fixed || (__builtin_isfinite(change) && change <= P.max_phase_change)) ? 0 : 1;
}

// propose_phi's bookkeeping (the one thread that writes the controller):
// the change for StepInfo, then a projection failure ends the step, a
// failed increment test decides the attempt, and otherwise the solve runs.
__device__ __forceinline__ void propose_record(DevScal* S, int verdict, double change) {
    S->runs[RUN_PROPOSE]++;
    S->change = change;
    if (verdict == 2) {
        S->status = ST_PROJECTION;
        S->phase = PH_IDLE;
        return;
    }
    S->live = verdict == 0 ? 1 : 0;
    S->phase = verdict == 0 ? PH_SOLVE : PH_DECIDE;
}

// The propose launch's result on its own (the single propose_phi entry; in
// a round prepare's finisher records it).
__global__ __launch_bounds__(TPB) void k_propose_fin(Bufs B) {
    const int d = blockIdx.x;
    const KP P = B.kp[d];
    DevScal* S = B.S + d;
    if (S->phase != PH_PROPOSE || S->need_state || threadIdx.x != 0) return;
    double change;
    propose_record(S, propose_verdict(P, S, change), change);
}

// Storage type of the Jacobi diagonal and the PCG vectors p and z: double in
// the plain form; float in the mixed form, where every value is used
// exactly as stored (promoted back to double).  These only steer the
// iteration -- the diagonal is the preconditioner, p and z the search
// direction, and x += alpha p, r -= alpha A p use the same stored p -- so the
// mixed solve reaches the same FP64 residual tolerance on the same FP64
// operator.  The operator's coefficients stay FP64: rounding chi moves
// mu = mu_ref + (c - intercept)/chi by its relative error times
// |mu - mu_ref| (1e-4 J/mol at FP32), and rounding K changes the face terms,
// about 0.4 of the diagonal at production dt, which accumulated to
// 3e-2 J/mol over 1000 steps.  x, r, Ap, rhs, the intercept and every dot
// product stay FP64.
template <bool MX>
using Coef = typename std::conditional<MX, float, double>::type;

template <bool MX>
__device__ __forceinline__ double stored(double v) {
    return static_cast<double>(static_cast<Coef<MX>>(v));
}

// Mixture, diffusivity as stored and mu - mu_ref of one cell of the trial
// state.
struct Prep {
    Mix m;
    double k;
    double x;
    bool ok;
};

// P2: implicit_diffusion setup (:1013-1020) fused with the initial
// residual, Jacobi diagonal and preconditioned residual (:1036-1045).
// Every cell computes its own mixture; the diffusivity of the neighbours
// comes from the column registers (south/north) or from the trial phases of
// the west/east cell (the same expression, hence the same bits).  Sums
// rhs.rhs, r.z, r.r and flags an invalid mixture; the finisher forms the
// tolerance and evaluates the loop head for iteration 0.  Neither rhs nor
// the initial x = mu_old - mu_ref is stored: the kernels that need them
// form them again from c_old, the trial phases and mu_old, with the same
// expressions.  z = r/diag is the first search direction p (:1042) and is
// stored as such, in the p buffer that cg_apply's first launch reads.
template <int TY, bool MX>
__global__ __launch_bounds__(TPB) void k_prepare(Bufs B, int slot) {
    using T = Coef<MX>;
    const Where w = where(B, -1);
    const KP P = B.kp[w.d];
    DevScal* S = B.S + w.d;
    // right after this round's propose launch: every block reads its
    // verdict, and the solve starts only if the attempt goes on
    const bool given = S->phase == PH_PREPARE;
    if (!(given || (S->phase == PH_PROPOSE && !S->need_state)) || P.phase_only) return;
    double change;
    if (!given && propose_verdict(P, S, change) != 0) return;
    record(B, slot, w.lb, 1);
    const int n = P.n;
    const int cur = S->cur;
    const double dt = S->dt;
    const double* __restrict__ phi_t = B.phi[cur ^ 1] + 3 * P.off;
    const double* __restrict__ c_old = B.c[cur] + P.off;
    const double* __restrict__ mu_old = B.mu[cur] + P.off;
    double* __restrict__ chi = B.chi + P.off;
    double* __restrict__ K = B.K + P.off;
    T* __restrict__ diag = reinterpret_cast<T*>(B.diag) + P.off;
    double* __restrict__ r = B.r + P.off;
    T* __restrict__ z = reinterpret_cast<T*>(S->pidx ? B.p0 : B.p1) + P.off;  // p = z
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    const Col col = column<TY>(P, w.lb, P.ny);
    // LDS tile: x = mu_old - mu_ref and K = mixture's K of the trial phases,
    // for the tile and its halo (each formed once per block).  The thread
    // that stages a column's centre rows keeps their whole mixture (chi and
    // c0; K is the same expression's value) in registers, so the trial
    // phases are read once per cell.
    __shared__ TileLds<TY> L;
    double mix_s[TY], mix_c0[TY];
    const int t = static_cast<int>(threadIdx.x);
    auto stage_column = [&](const int cc, int gx, const bool centre) {
        gx = gx < 0 ? gx + P.nx : (gx >= P.nx ? gx % P.nx : gx);
#pragma unroll
        for (int rw = 0; rw < TY + 2; ++rw) {
            int y = col.y0 - 1 + rw;
            y = y < 0 ? 0 : (y > P.ny - 1 ? P.ny - 1 : y);
            const int j = y * P.nx + gx;
            double p[3];
            load_phi(phi_t, n, j, p);
            L.a[rw][cc] = mu_old[j] - P.mu_ref;
            if (centre && rw >= 1 && rw <= TY) {
                Mix m;
                const bool ok = mixture(P, p, m);
                if (!ok && col.on && col.y0 + rw - 1 < col.y1) v[3] = 1.0;
                mix_s[rw - 1] = m.s;
                mix_c0[rw - 1] = m.c0;
                L.b[rw][cc] = m.K;
            } else {
                L.b[rw][cc] = mixture_K(P, p);
            }
        }
    };
    stage_column(t + 1, col.x, true);
    if (t == 0) stage_column(0, col.x - 1, false);
    if (t == TPB - 1) stage_column(TPB + 1, col.x + 1, false);
    __syncthreads();
    for_tile<TY>(P, col, [&](const int i, const int y, const int row, const int cc) {
        Mix m;
        m.s = mix_s[row - 1];
        m.c0 = mix_c0[row - 1];
        m.K = 0.0;  // not used: composition needs chi and c0
        const double ic = composition(P.mu_ref, m);
        const double b = ld_once(c_old + i) - ic;
        const double ch = m.s;
        const double kc = L.b[row][cc], xc = L.a[row][cc];
        st_once(chi + i, ch);
        st_once(K + i, kc);
        double wf[4];
        face_weights(P, y, kc, L.b[row][cc - 1], L.b[row][cc + 1], L.b[row - 1][cc], L.b[row + 1][cc], wf);
        const double ax =
            apply_vals(ch, dt, wf, xc, L.a[row][cc - 1], L.a[row][cc + 1], L.a[row - 1][cc], L.a[row + 1][cc]);
        const double d = stored<MX>(ch + dt * (wf[0] + wf[1] + wf[2] + wf[3]));
        const double rr = b - ax;
        const double zz = stored<MX>(rr / d);
        st_once(diag + i, static_cast<T>(d));
        st_once(r + i, rr);
        st_once(z + i, static_cast<T>(zz));
        v[0] += b * b;
        v[1] += rr * zz;
        v[2] += rr * rr;
    });
    block_partial<4, (1u << 6)>(v, B.part + P.poff, P.nb, w.lb);
}

// Records propose's result (propose_record) and, when the solve runs,
// finishes prepare's sums.
__global__ __launch_bounds__(TPB) void k_prepare_fin(Bufs B) {
    const int d = blockIdx.x;
    const KP P = B.kp[d];
    DevScal* S = B.S + d;
    const bool given = S->phase == PH_PREPARE;
    if (!(given || (S->phase == PH_PROPOSE && !S->need_state))) return;
    double change = 0;
    const int verdict = given ? 0 : propose_verdict(P, S, change);
    if (verdict != 0 || P.phase_only) {
        if (threadIdx.x == 0) propose_record(S, verdict, change);
        return;
    }
    double v[4];
    if (!final_reduce<4, (1u << 6)>(v, B.part + P.poff, P.nb)) return;
    if (given) S->phase = PH_SOLVE;
    else propose_record(S, verdict, change);
    S->runs[RUN_PREPARE]++;
    if (v[3] > 0) {
        S->status = ST_MIXTURE;
        S->phase = PH_IDLE;
        return;
    }
    S->rhs2 = v[0];
    S->rz = v[1];
    S->rr = v[2];
    S->tol = P.cg_atol + P.cg_rtol * __builtin_sqrt(v[0]);
    S->restart = 1;
    S->pad = 0;
    S->x_mu = 1;
    S->x_lag = 0;
    cg_head(P, S, 0);
}

// The forward solve's solution at cell j as the kernel after the last
// update sees it (DevScal::x_mu, x_lag): the x plane, or mu_old - mu_ref
// before the plane is first written (prepare's expression), plus the last
// step alpha p where it is still to be applied (cg_update's expression).
template <bool MU, bool LAG, class T>
__device__ __forceinline__ double x_at(const double* __restrict__ x, const double* __restrict__ mu_old, double mu_ref,
                                       const T* __restrict__ p, double alpha, int j) {
    const double b = MU ? mu_old[j] - mu_ref : x[j];
    return LAG ? b + alpha * static_cast<double>(p[j]) : b;
}

// f(MU, LAG) with the two flags as compile-time constants (uniform per
// domain, so each kernel runs one branch-free form).
template <class F>
__device__ __forceinline__ void with_x_form(int mu, int lag, F&& f) {
    if (mu) {
        if (lag) f(std::true_type{}, std::true_type{});
        else f(std::true_type{}, std::false_type{});
    } else {
        if (lag) f(std::false_type{}, std::true_type{});
        else f(std::false_type{}, std::false_type{});
    }
}

// PCG kernel A.  In state CG_ITER: the search direction of this iteration
// p = z + beta p_old (:1079), kept in the other p buffer (after a
// (re)start, p = z :1042/:1062 is already there: prepare or the
// true-residual pass stored it), then Ap = A p (:1068) and p.Ap; the finisher checks
// for breakdown (:1070) and forms the step length (:1071).  In the forward
// solve it also applies the previous step's x += alpha p_old (:1074), which
// cg_update leaves to it: this kernel reads p_old anyway, so x costs its
// read and write here instead of x and p reads and the x write there.  In
// state CG_TRUE: the true residual r = rhs - A x with z = r/diag
// (:1053-1062), rhs formed from c_old and the trial phases as prepare forms
// it (the adjoint's solve stores its rhs); the finisher accepts (:1056) or
// restarts from it (:1060-1065).
template <int TY, bool MX>
__global__ __launch_bounds__(TPB) void k_cg_a(Bufs B, int slot) {
    using T = Coef<MX>;
    const Where w = where(B, -1);
    const KP P = B.kp[w.d];
    DevScal* S = B.S + w.d;
    if (S->phase != PH_SOLVE) return;
    const int state = S->cg_state;
    const double dt = S->dt;
    const double* __restrict__ chi = B.chi + P.off;
    const double* __restrict__ K = B.K + P.off;
    T* __restrict__ z = reinterpret_cast<T*>(B.z) + P.off;
    double* __restrict__ x = B.x + P.off;
    const double* __restrict__ mu_old = B.mu[S->cur] + P.off;
    const int pi = S->pidx;
    const T* __restrict__ po = reinterpret_cast<const T*>(pi ? B.p1 : B.p0) + P.off;
    const double alpha = S->alpha;  // the previous step's (its finisher has not run yet)
    T* __restrict__ pn = reinterpret_cast<T*>(pi ? B.p0 : B.p1) + P.off;
    const Col col = column<TY>(P, w.lb, P.ny);
    __shared__ TileLds<TY> L;  // a: the stencil variable, b: K
    if (state == CG_ITER) {
        record(B, slot, w.lb, S->x_lag ? 1 : 3);
        double* __restrict__ Ap = B.Ap + P.off;
        const double beta = S->beta;
        double v[1] = {0.0};
        // the direction as it is stored; A p is formed from exactly these
        // values.  The restart flag is uniform: one staging loop per value
        // keeps the p_old loads free of branches.  The thread keeps p_old of
        // its own cells for the step x += alpha p_old.
        double pc[TY];
        auto stage = [&](auto fresh) {
            stage_tile<TY>(P, col, [&](const int r, const int cc, const int j, auto own) {
                if constexpr (decltype(fresh)::value) {
                    L.a[r][cc] = static_cast<double>(pn[j]);
                } else {
                    const double pv = static_cast<double>(po[j]);
                    if constexpr (decltype(own)::value)
                        if (r >= 1 && r <= TY) pc[r - 1] = pv;
                    L.a[r][cc] = stored<MX>(z[j] + beta * pv);
                }
                L.b[r][cc] = K[j];
            });
        };
        auto body = [&](auto fresh, auto mu, auto lag) {
            for_tile<TY>(P, col, [&](const int i, const int y, const int r, const int cc) {
                double wf[4];
                face_weights(P, y, L.b[r][cc], L.b[r][cc - 1], L.b[r][cc + 1], L.b[r - 1][cc], L.b[r + 1][cc], wf);
                const double c = L.a[r][cc];
                const double a = apply_vals(ld_once(chi + i), dt, wf, c, L.a[r][cc - 1], L.a[r][cc + 1],
                                            L.a[r - 1][cc], L.a[r + 1][cc]);
                if constexpr (!decltype(fresh)::value) st_once(pn + i, static_cast<T>(c));
                st_once(Ap + i, a);
                if constexpr (decltype(lag)::value) {
                    // x + alpha p_old (cg_update's expression), p_old as stored
                    const double pv = decltype(fresh)::value ? static_cast<double>(po[i]) : pc[r - 1];
                    const double b = decltype(mu)::value ? mu_old[i] - P.mu_ref : ld_once(x + i);
                    st_once(x + i, b + alpha * pv);
                }
                v[0] += c * a;
            });
        };
        auto go = [&](auto fresh) {
            stage(fresh);
            __syncthreads();
            with_x_form(S->x_mu, S->x_lag, [&](auto mu, auto lag) { body(fresh, mu, lag); });
        };
        if (S->restart != 0) go(std::true_type{});
        else go(std::false_type{});
        block_partial<1, 0u>(v, B.part + P.poff, P.nb, w.lb);
    } else if (state == CG_TRUE) {
        record(B, slot, w.lb, 2);
        const double* __restrict__ rhs = B.rhs + P.off;
        const double* __restrict__ c_old = B.c[S->cur] + P.off;
        const double* __restrict__ phi_t = B.phi[S->cur ^ 1] + 3 * P.off;
        const T* __restrict__ diag = reinterpret_cast<const T*>(B.diag) + P.off;
        double* __restrict__ r = B.r + P.off;
        const bool alone = S->cg_alone != 0;
        double v[2] = {0.0, 0.0};
        with_x_form(S->x_mu, S->x_lag, [&](auto mu, auto lag) {
            stage_tile<TY>(P, col, [&](const int rr, const int cc, const int j, auto) {
                L.a[rr][cc] = x_at<decltype(mu)::value, decltype(lag)::value>(x, mu_old, P.mu_ref, po, alpha, j);
                L.b[rr][cc] = K[j];
            });
        });
        __syncthreads();
        for_tile<TY>(P, col, [&](const int i, const int y, const int rw, const int cc) {
            double wf[4];
            face_weights(P, y, L.b[rw][cc], L.b[rw][cc - 1], L.b[rw][cc + 1], L.b[rw - 1][cc], L.b[rw + 1][cc], wf);
            double b;
            if (alone) {
                b = ld_once(rhs + i);
            } else {
                double q[3];
                load_phi(phi_t, P.n, i, q);
                Mix m;
                (void)mixture(P, q, m);
                b = ld_once(c_old + i) - composition(P.mu_ref, m);
            }
            const double rr = b - apply_vals(ld_once(chi + i), dt, wf, L.a[rw][cc], L.a[rw][cc - 1], L.a[rw][cc + 1],
                                             L.a[rw - 1][cc], L.a[rw + 1][cc]);
            const double zz = stored<MX>(rr / static_cast<double>(ld_once(diag + i)));
            st_once(r + i, rr);
            st_once(pn + i, static_cast<T>(zz));  // the restart's direction p = z (:1062)
            v[0] += rr * rr;
            v[1] += rr * zz;
        });
        block_partial<2, 0u>(v, B.part + P.poff, P.nb, w.lb);
    }
}

__global__ __launch_bounds__(TPB) void k_cg_a_fin(Bufs B) {
    const int d = blockIdx.x;
    const KP P = B.kp[d];
    DevScal* S = B.S + d;
    if (S->phase != PH_SOLVE) return;
    const int state = S->cg_state;
    if (state == CG_ITER) {
        double v[1];
        if (!final_reduce<1, 0u>(v, B.part + P.poff, P.nb)) return;
        if (S->x_lag) {
            S->runs[RUN_CG_APPLY]++;
            S->x_lag = 0;
            S->x_mu = 0;
        } else {
            S->runs[RUN_CG_FRESH]++;
        }
        const double pAp = v[0], rz = S->rz;
        S->pAp = pAp;
        if (!(pAp > 0) || !__builtin_isfinite(pAp) || !__builtin_isfinite(rz)) {
            S->cg_state = CG_FAIL;
            S->phase = PH_DECIDE;
        } else {
            S->alpha = rz / pAp;
            S->pidx ^= 1;
            S->pad = 1;
        }
    } else if (state == CG_TRUE) {
        double v[2];
        if (!final_reduce<2, 0u>(v, B.part + P.poff, P.nb)) return;
        S->runs[RUN_TRUE_RESIDUAL]++;
        S->rr = v[0];
        const double res = __builtin_sqrt(v[0]);
        S->cg_res = res;
        if (res <= S->tol * 1.05) {
            S->cg_state = CG_OK;
        } else if (S->cg_it >= P.cg_max) {
            // the restart's `continue` (:1065) leaves the loop after its
            // last iteration: not converged, info.it = cg_max
            S->cg_state = CG_FAIL;
            S->phase = PH_DECIDE;
        } else {
            S->rz = v[1];
            S->restart = 1;
            cg_head(P, S, S->cg_it + 1);
        }
    }
}

// PCG kernel B, after a step of kernel A: r -= alpha Ap; z = r/diag
// (:1074-1076); sums r.z and r.r; the finisher forms beta (:1077) and
// evaluates the loop head of the next iteration.  x += alpha p (:1074) is
// left to the next kernel that reads x or p (DevScal::x_lag), except in the
// adjoint's solve, whose result is the x plane itself.
// Threads per column of cg_update.  Its cells are independent, so two
// threads share a column's rows and twice as many loads are in flight; the
// fold stays the column's sequential one: the thread of the upper rows hands
// its products to the thread of the lower rows through LDS, which adds them
// after its own, in row order.
template <int TY>
__host__ __device__ constexpr int cgb_split() { return TY >= 2 ? 2 : 1; }

template <int TY, bool MX>
__global__ __launch_bounds__(TPB * cgb_split<TY>()) void k_cg_b(Bufs B, int slot) {
    constexpr int SPLIT = cgb_split<TY>(), RH = TY / SPLIT;  // rows per thread
    using T = Coef<MX>;
    const Where w = where(B, -1);
    const KP P = B.kp[w.d];
    DevScal* S = B.S + w.d;
    if (!(S->phase == PH_SOLVE && S->cg_state == CG_ITER && S->pad == 1)) return;
    record(B, slot, w.lb, 1);
    const T* __restrict__ p = reinterpret_cast<const T*>(S->pidx ? B.p1 : B.p0) + P.off;
    double* __restrict__ x = B.x + P.off;
    double* __restrict__ r = B.r + P.off;
    T* __restrict__ z = reinterpret_cast<T*>(B.z) + P.off;
    const double* __restrict__ Ap = B.Ap + P.off;
    const T* __restrict__ diag = reinterpret_cast<const T*>(B.diag) + P.off;
    const double alpha = S->alpha;
    const bool upd_x = S->cg_alone != 0;
    const int t = static_cast<int>(threadIdx.x) % TPB, h = static_cast<int>(threadIdx.x) / TPB;
    const Col col = column<TY>(P, w.lb, P.ny, t);
    double prod[RH][2];
    int cnt = 0;
    if (col.on) {
        int i = col.i0 + h * RH * P.nx;
#pragma unroll
        for (int k = 0; k < RH; ++k) {
            const int y = col.y0 + h * RH + k;
            if (y >= col.y1) break;
            // nontemporal throughout: measured against cg_apply run alternately
            // with it, the pair is 9 % faster than with plain accesses (plain
            // stores leave dirty lines that the next cg_apply pays for)
            if (upd_x) st_once(x + i, ld_once(x + i) + alpha * static_cast<double>(ld_once(p + i)));
            const double rr = ld_once(r + i) - alpha * ld_once(Ap + i);
            const double zz = stored<MX>(rr / static_cast<double>(ld_once(diag + i)));
            st_once(r + i, rr);
            st_once(z + i, static_cast<T>(zz));
            prod[k][0] = rr * zz;
            prod[k][1] = rr * rr;
            ++cnt;
            i += P.nx;
        }
    }
    double v[2] = {0.0, 0.0};
    if (h == 0) {
#pragma unroll
        for (int k = 0; k < RH; ++k) {
            if (k >= cnt) break;
            v[0] += prod[k][0];
            v[1] += prod[k][1];
        }
    }
    if constexpr (SPLIT == 2) {
        __shared__ double hand[RH][2][TPB];
        __shared__ int hcnt[TPB];
        if (h == 1) {
#pragma unroll
            for (int k = 0; k < RH; ++k) {
                if (k >= cnt) break;
                hand[k][0][t] = prod[k][0];
                hand[k][1][t] = prod[k][1];
            }
            hcnt[t] = cnt;
        }
        __syncthreads();
        if (h == 0) {
            const int m = hcnt[t];
#pragma unroll
            for (int k = 0; k < RH; ++k) {
                if (k >= m) break;
                v[0] += hand[k][0][t];
                v[1] += hand[k][1][t];
            }
        }
    }
    block_partial<2, 0u>(v, B.part + P.poff, P.nb, w.lb);
}

__global__ __launch_bounds__(TPB) void k_cg_b_fin(Bufs B) {
    const int d = blockIdx.x;
    const KP P = B.kp[d];
    DevScal* S = B.S + d;
    if (!(S->phase == PH_SOLVE && S->cg_state == CG_ITER && S->pad == 1)) return;
    double v[2];
    if (!final_reduce<2, 0u>(v, B.part + P.poff, P.nb)) return;
    S->runs[RUN_CG_UPDATE]++;
    const double next = v[0];
    S->beta = next / S->rz;
    S->rz = next;
    S->rr = v[1];
    S->restart = 0;
    S->pad = 0;
    S->x_lag = S->cg_alone ? 0 : 1;
    cg_head(P, S, S->cg_it + 1);
}

// Solution, diffusivity and trial phases of one cell.
struct Fin {
    double x, k, p[3];
};

// F + A: flux-form carbon update, closure, mu reconstruction (:1082-1093),
// fused with the acceptance sweep on the trial state: valid_state (:1096),
// total_carbon (:784) and reduced_energy (:793).  In state CG_CHECK it also
// forms the true residual r = rhs - A x of the converged solve (:1053-1055)
// and sums r.r: every operand is already here (x and K in the tile, chi and
// the intercept from the trial phases, rhs = c_old - intercept as prepare
// forms it), so r is the true-residual pass's value bit for bit and the sum
// keeps its fold order (rows in order per thread, then the same tree).  The
// finisher keeps the update when the check passes (:1056) and otherwise
// hands the solve back to the true-residual pass, which recomputes r and z
// and restarts (:1060-1065); the discarded update is rewritten later.  x is
// staged as the last update left it: with the last step alpha p applied
// here (cg_update's expression) and, after a solve of at most one step,
// from mu_old (prepare's expression).
template <int TY, bool MX>
__global__ __launch_bounds__(TPB) void k_finalize(Bufs B, int slot) {
    using T = Coef<MX>;
    const Where w = where(B, -1);
    const KP P = B.kp[w.d];
    DevScal* S = B.S + w.d;
    if (!(S->phase == PH_SOLVE && (S->cg_state == CG_OK || S->cg_state == CG_CHECK))) return;
    record(B, slot, w.lb, 1);
    const int n = P.n;
    const int cur = S->cur;
    const double dt = S->dt;
    const double* __restrict__ phi_t = B.phi[cur ^ 1] + 3 * P.off;
    const double* __restrict__ K = B.K + P.off;
    const double* __restrict__ x = B.x + P.off;
    const double* __restrict__ mu_old = B.mu[cur] + P.off;
    const T* __restrict__ p_last = reinterpret_cast<const T*>(S->pidx ? B.p1 : B.p0) + P.off;
    const double alpha = S->alpha;
    const double* __restrict__ c_old = B.c[cur] + P.off;
    double* __restrict__ c_new = B.c[cur ^ 1] + P.off;
    double* __restrict__ mu_new = B.mu[cur ^ 1] + P.off;
    double v[5] = {0.0, 1.0, 0.0, 0.0, 0.0};
    bool good = true, mix_ok = true;
    const Col col = column<TY>(P, w.lb, P.ny);
    // LDS tile: x and K of the tile and its halo; the trial phases march
    // down the column in registers (the energy needs east and north)
    __shared__ TileLds<TY> L;
    with_x_form(S->x_mu, S->x_lag, [&](auto mu, auto lag) {
        stage_tile<TY>(P, col, [&](const int r, const int cc, const int j, auto) {
            L.a[r][cc] = x_at<decltype(mu)::value, decltype(lag)::value>(x, mu_old, P.mu_ref, p_last, alpha, j);
            L.b[r][cc] = K[j];
        });
    });
    __syncthreads();
    const int cc = static_cast<int>(threadIdx.x) + 1;
    march<TY, Phases>(
        P, col,
        [&](const int j) {
            Phases q;
            load_phi(phi_t, n, j, q.p);
            return q;
        },
        [&](const int i, const int y, const Phases&, const Phases& c, const Phases& nn) {
            const int r = y - col.y0 + 1;
            double wf[4];
            face_weights(P, y, L.b[r][cc], L.b[r][cc - 1], L.b[r][cc + 1], L.b[r - 1][cc], L.b[r + 1][cc], wf);
            const double xi = L.a[r][cc];
            const double change = dt * (wf[0] * (L.a[r][cc - 1] - xi) + wf[1] * (L.a[r][cc + 1] - xi) +
                                        wf[2] * (L.a[r - 1][cc] - xi) + wf[3] * (L.a[r + 1][cc] - xi));
            const double co = ld_once(c_old + i);
            const double cn = co + change;
            // the intercept and susceptibility prepare formed from the same
            // trial phases: the same expressions, hence the same values
            Mix m;
            if (!mixture(P, c.p, m)) mix_ok = false;
            const double ic = composition(P.mu_ref, m), ch = m.s;
            const double rt = (co - ic) - apply_vals(ch, dt, wf, xi, L.a[r][cc - 1], L.a[r][cc + 1], L.a[r - 1][cc],
                                                     L.a[r + 1][cc]);
            v[4] += rt * rt;
            v[0] = dmaxd(v[0], __builtin_fabs(cn - (ic + ch * xi)));
            const double mn = P.mu_ref + (cn - ic) / ch;
            st_once(c_new + i, cn);
            st_once(mu_new + i, mn);
            double qe[3];
            load_phi(phi_t, n, i + col.de, qe);
            double e;
            accept_vals(P, c.p, qe, nn.p, y != P.ny - 1, cn, mn, good, e, mix_ok, &m);
            v[2] += cn;
            v[3] += e;
        });
    v[1] = good ? 1.0 : 0.0;
    block_partial<5, (1u | (2u << 2))>(v, B.part + P.poff, P.nb, w.lb);
}

// finalize's finisher for domain d (one 256-thread block; k_finalize_fin)
__device__ __forceinline__ void finalize_fin_domain(const Bufs& B, int d) {
    const KP P = B.kp[d];
    DevScal* S = B.S + d;
    if (!(S->phase == PH_SOLVE && (S->cg_state == CG_OK || S->cg_state == CG_CHECK))) return;
    double v[5];
    if (!final_reduce<5, (1u | (2u << 2))>(v, B.part + P.poff, P.nb)) return;
    S->runs[RUN_FINALIZE]++;
    if (S->x_lag) S->runs[RUN_FIN_LAG]++;
    if (S->cg_state == CG_CHECK) {
        S->rr = v[4];
        const double res = __builtin_sqrt(v[4]);
        S->cg_res = res;
        if (!(res <= S->tol * 1.05)) {
            S->cg_state = CG_TRUE;
            return;
        }
        S->cg_state = CG_OK;
    }
    S->closure = v[0];
    S->valid = v[1];
    S->mass = v[2];
    S->energy = v[3];
    S->phase = PH_DECIDE;
}

// Phase-only acceptance (:1134-1136): carbon = composition(mu, mixture(phi')),
// mu unchanged; then valid_state and reduced_energy on the trial state.
template <int TY>
__global__ __launch_bounds__(TPB) void k_phase_only(Bufs B, int slot) {
    const Where w = where(B, -1);
    const KP P = B.kp[w.d];
    DevScal* S = B.S + w.d;
    if (!(S->phase == PH_SOLVE && P.phase_only)) return;
    record(B, slot, w.lb, 1);
    const int cur = S->cur;
    const double* __restrict__ phi_t = B.phi[cur ^ 1] + 3 * P.off;
    const double* __restrict__ mu_old = B.mu[cur] + P.off;
    double* __restrict__ c_new = B.c[cur ^ 1] + P.off;
    double* __restrict__ mu_new = B.mu[cur ^ 1] + P.off;
    double v[5] = {0.0, 1.0, 0.0, 0.0, 0.0};
    bool good = true, mix_ok = true;
    const Col col = column<TY>(P, w.lb, P.ny);
    for_column<TY>(P, col, [&](const int i, const int y) {
        double p[3];
        load_phi(phi_t, P.n, i, p);
        Mix m;
        if (!mixture(P, p, m)) mix_ok = false;
        const double mu = mu_old[i];
        const double cn = composition(mu, m);
        c_new[i] = cn;
        mu_new[i] = mu;
        double e;
        accept_cell(P, phi_t, i, nbr_of(P, col, i, y), cn, mu, good, e, mix_ok);
        v[2] += cn;
        v[3] += e;
    });
    v[1] = good ? 1.0 : 0.0;
    v[4] = mix_ok ? 0.0 : 1.0;
    block_partial<5, (1u | (2u << 2) | (1u << 8))>(v, B.part + P.poff, P.nb, w.lb);
}

// the phase-only acceptance's finisher for domain d (k_phase_only_fin)
__device__ __forceinline__ void phase_only_fin_domain(const Bufs& B, int d) {
    const KP P = B.kp[d];
    DevScal* S = B.S + d;
    if (!(S->phase == PH_SOLVE && P.phase_only)) return;
    double v[5];
    if (!final_reduce<5, (1u | (2u << 2) | (1u << 8))>(v, B.part + P.poff, P.nb)) return;
    S->runs[RUN_FINALIZE]++;
    if (v[4] > 0) {
        S->status = ST_MIXTURE;
        S->phase = PH_IDLE;
        return;
    }
    S->closure = 0;
    S->valid = v[1];
    S->mass = v[2];
    S->energy = v[3];
    S->phase = PH_DECIDE;
}

// Acceptance tests and the dt controller of advance (:1131-1173), for
// domain d.
__device__ __forceinline__ void decide_domain(const Bufs& B, int d) {
    const KP& P = B.kp[d];
    DevScal* S = B.S + d;
    if (S->phase != PH_DECIDE) return;
    if (S->fixed) {
        // a fixed-dt step: taken whenever the solve converged, no acceptance
        // tests, no dt control, the cached e0/mass0 dropped
        if (S->cg_state != CG_OK) {
            S->status = ST_FIXED_SOLVE;
            S->phase = PH_IDLE;
            return;
        }
        S->cur ^= 1;
        S->time = S->time + S->dt;
        S->step = S->step + 1;
        S->e0_valid = 0;
        S->m0_valid = 0;
        if (S->cg_it > S->cg_peak) S->cg_peak = S->cg_it;
        S->taken++;
        S->phase = S->taken >= S->target_steps ? PH_IDLE : PH_START;
        return;
    }
    bool good = S->live != 0;
    int why = 0;  // phase increment
    double residual = 0, closure = 0;
    int it = 0;
    if (good && !P.phase_only) {
        it = S->cg_it;
        residual = S->cg_res;
        S->cg_all += it;
        closure = S->cg_state == CG_OK ? S->closure : 0;
        good = S->cg_state == CG_OK && closure <= P.closure_tol;
        why = 1;  // linear/closure
    }
    const double mass = S->mass * P.dx * P.dy;
    const double energy = S->energy * P.dx * P.dy;
    if (good) {
        good = S->valid > 0.5;
        why = 2;  // composition/phase bounds
    }
    if (good && !P.phase_only) {
        good = __builtin_fabs(mass - S->m0) <= 1e-11 * dmaxd(__builtin_fabs(S->m0), P.mass_floor);
        why = 3;  // mass conservation
    }
    double ech = 0;
    if (good && P.check_energy) {
        ech = energy - S->e0;
        const double tol = P.energy_atol + P.energy_rtol * dmaxd(__builtin_fabs(S->e0), P.energy_scale);
        good = __builtin_isfinite(energy) && energy <= S->e0 + tol;
        why = 4;  // free-energy increase
    }
    S->attempts++;
    if (good) {
        const double dt = S->dt;
        S->cur ^= 1;
        S->time = S->time + dt;
        S->step = S->step + 1;
        S->rejected = S->rejected + S->attempt;
        S->dt_state = dmind(P.dt_max, dt * 1.1);
        if (P.check_energy) {
            S->e0 = energy;
            S->e0_valid = 1;
        }
        if (!P.phase_only) {
            S->m0 = mass;
            S->m0_valid = 1;
        }
        S->i_dt = dt;
        S->i_change = S->change;
        S->i_residual = residual;
        S->i_closure = closure;
        S->i_energy_change = ech;
        S->i_cg = it;
        S->i_retries = S->attempt;
        S->cg_total += it;
        if (it > S->cg_peak) S->cg_peak = it;
        S->taken++;
        S->why = -1;
        S->phase = (S->taken >= S->target_steps || !(S->time < S->stop_time - S->stop_eps)) ? PH_IDLE : PH_START;
        return;
    }
    S->rejected_attempts++;
    S->why = why;
    S->dt = S->dt * .5;
    if (S->dt < P.dt_min) {
        S->status = ST_DT_MIN;
        S->phase = PH_IDLE;
    } else if (S->attempt >= P.max_retries) {
        S->status = ST_RETRY_LIMIT;
        S->phase = PH_IDLE;
    } else {
        S->attempt++;
        S->phase = PH_PROPOSE;
    }
}

// The end of a round for domain d, the turn: the decision on the attempt,
// then the start of the next step or attempt.
__device__ __forceinline__ void turn_domain(const Bufs& B, int d) {
    decide_domain(B, d);
    begin_domain(B, d);
}

// The finishers of the acceptance passes, one block per domain.  TURN: the
// finisher is the round's last launch and also takes its domain's turn, so
// the turn needs no launch of its own (each launch costs ~4.5 us however
// little it does).  Every domain's block takes the turn, whether or not its
// finisher had work this round.  The barrier orders the turn's writes to the
// controller after every thread's reading of it.
template <bool TURN>
__global__ __launch_bounds__(TPB) void k_finalize_fin(Bufs B) {
    const int d = blockIdx.x;
    finalize_fin_domain(B, d);
    if (TURN) {
        __syncthreads();
        if (threadIdx.x == 0) turn_domain(B, d);
    }
}

template <bool TURN>
__global__ __launch_bounds__(TPB) void k_phase_only_fin(Bufs B) {
    const int d = blockIdx.x;
    phase_only_fin_domain(B, d);
    if (TURN) {
        __syncthreads();
        if (threadIdx.x == 0) turn_domain(B, d);
    }
}

// Fixed-dt runs that keep every state (the adjoint's segment recompute):
// after each taken step, the current phi (3 planes), c and mu of domain 0
// are copied to the next slot of fixed_save.  Grid-stride over 5n values.
__global__ __launch_bounds__(TPB) void k_fixed_save(Bufs B) {
    const KP P = B.kp[0];
    DevScal* S = B.S;
    if (!S->fixed || !S->fixed_save || S->fixed_saved >= S->taken) return;
    const long long n = P.n, k = S->taken;
    double* dst = S->fixed_save + (k - 1) * 5 * n;
    const int cur = S->cur;
    const double* phi = B.phi[cur] + 3 * P.off;
    const double* c = B.c[cur] + P.off;
    const double* mu = B.mu[cur] + P.off;
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < 5 * n; i += stride) {
        const double v = i < 3 * n ? phi[i] : (i < 4 * n ? c[i - 3 * n] : mu[i - 4 * n]);
        st_once(dst + i, v);
    }
}

// Marks the step copied by k_fixed_save (a separate launch, so no block of
// the copy can see the mark before the copy is done).
__global__ void k_fixed_saved(Bufs B) {
    DevScal* S = B.S;
    if (threadIdx.x == 0 && S->fixed && S->fixed_save) S->fixed_saved = static_cast<int>(S->taken);
}

__global__ void k_clear_kinds(int* kinds, int n) {
    const int i = threadIdx.x;
    if (i < n) kinds[i] = 0;
}

// ---------------------------------------------------------------------------
// Direct (single-domain) kernels: state reductions, diagnostics, window.
// ---------------------------------------------------------------------------

// valid_state, total_carbon and reduced_energy of domain d's current state.
template <int TY>
__global__ __launch_bounds__(TPB) void k_state_direct(Bufs B, int d) {
    const KP P = B.kp[d];
    DevScal* S = B.S + d;
    const int lb = blockIdx.x;
    const int cur = S->cur;
    const double* phi = B.phi[cur] + 3 * P.off;
    const double* c = B.c[cur] + P.off;
    const double* mu = B.mu[cur] + P.off;
    double v[4] = {1.0, 0.0, 0.0, 0.0};
    bool good = true, mix_ok = true;
    const Col col = column<TY>(P, lb, P.ny);
    for_column<TY>(P, col, [&](const int i, const int y) {
        state_cell(P, phi, c, mu, i, nbr_of(P, col, i, y), good, v[1], v[2], mix_ok);
    });
    v[0] = good ? 1.0 : 0.0;
    v[3] = mix_ok ? 0.0 : 1.0;
    block_partial<4, (2u | (1u << 6))>(v, B.part + P.poff, P.nb, lb);
}

__global__ __launch_bounds__(TPB) void k_state_direct_fin(Bufs B, int d) {
    const KP P = B.kp[d];
    DevScal* S = B.S + d;
    double v[4];
    if (!final_reduce<4, (2u | (1u << 6))>(v, B.part + P.poff, P.nb)) return;
    S->e_valid = v[0];
    S->e_mass = v[1];
    S->e_energy = v[2];
    S->err = v[3] > 0 ? (S->err | ERRF_MIXTURE_STATE) : (S->err & ~ERRF_MIXTURE_STATE);
}

// Snapshot reductions (:2267-2282): min/max of carbon and mu, max phase-sum
// error, carbon sum, energy.
constexpr unsigned DIAG_OPS = 2u | (1u << 2) | (2u << 4) | (1u << 6) | (1u << 8) | (1u << 14);
template <int TY>
__global__ __launch_bounds__(TPB) void k_diagnostics(Bufs B, int d) {
    const KP P = B.kp[d];
    DevScal* S = B.S + d;
    const int lb = blockIdx.x;
    const int cur = S->cur;
    const double* phi = B.phi[cur] + 3 * P.off;
    const double* c = B.c[cur] + P.off;
    const double* mu = B.mu[cur] + P.off;
    double v[8] = {1e300, -1e300, 1e300, -1e300, 0.0, 0.0, 0.0, 0.0};
    bool good = true, mix_ok = true;
    const Col col = column<TY>(P, lb, P.ny);
    for_column<TY>(P, col, [&](const int i, const int y) {
        const double ci = c[i], mi = mu[i];
        v[0] = dmind(v[0], ci);
        v[1] = dmaxd(v[1], ci);
        v[2] = dmind(v[2], mi);
        v[3] = dmaxd(v[3], mi);
        double p[3];
        load_phi(phi, P.n, i, p);
        v[4] = dmaxd(v[4], __builtin_fabs(p[0] + p[1] + p[2] - 1));
        double e;
        accept_cell(P, phi, i, nbr_of(P, col, i, y), ci, mi, good, e, mix_ok);
        v[5] += ci;
        v[6] += e;
    });
    v[7] = mix_ok ? 0.0 : 1.0;
    block_partial<8, DIAG_OPS>(v, B.part + P.poff, P.nb, lb);
}

__global__ __launch_bounds__(TPB) void k_diagnostics_fin(Bufs B, int d) {
    const KP P = B.kp[d];
    DevScal* S = B.S + d;
    double v[8];
    if (!final_reduce<8, DIAG_OPS>(v, B.part + P.poff, P.nb)) return;
    S->dmin_c = v[0];
    S->dmax_c = v[1];
    S->dmin_mu = v[2];
    S->dmax_mu = v[3];
    S->dsumerr = v[4];
    S->dmass = v[5];
    S->denergy = v[6];
    S->err = v[7] > 0 ? (S->err | ERRF_MIXTURE_STATE) : (S->err & ~ERRF_MIXTURE_STATE);
}

// Moving-window front (:2057-2062): max over interior faces of the gamma=1/2
// crossing height.
template <int TY>
__global__ __launch_bounds__(TPB) void k_front(Bufs B, int d) {
    const KP P = B.kp[d];
    DevScal* S = B.S + d;
    const int lb = blockIdx.x;
    const double* g = B.phi[S->cur] + 3 * P.off + 2 * static_cast<long long>(P.n);
    double v[1] = {-1.0};
    const Col col = column<TY>(P, lb, P.ny);
    for_column<TY>(P, col, [&](const int i, const int y) {
        if (y < P.ny - 1) {
            const double a = 1 - 2 * g[i], b = 1 - 2 * g[i + P.nx];
            if (a >= 0 && b < 0) v[0] = dmaxd(v[0], ((y + 1) - .5 + a / (a - b)) * P.dy);
        }
    });
    block_partial<1, 1u>(v, B.part + P.poff, P.nb, lb);
}

__global__ __launch_bounds__(TPB) void k_front_fin(Bufs B, int d) {
    const KP P = B.kp[d];
    double v[1];
    if (final_reduce<1, 1u>(v, B.part + P.poff, P.nb)) B.S[d].front = v[0];
}

// Max gamma and removed carbon (c*dx)*dy over the first `rows` rows (:2028,
// :2067), in the tile order of those rows.  Launched with
// nbs = tiles_x * ceil(rows/TY) blocks.
template <int TY>
__global__ __launch_bounds__(TPB) void k_strip(Bufs B, int d, int rows, int nbs) {
    const KP P = B.kp[d];
    DevScal* S = B.S + d;
    const int lb = blockIdx.x;
    const int cur = S->cur;
    const double* g = B.phi[cur] + 3 * P.off + 2 * static_cast<long long>(P.n);
    const double* c = B.c[cur] + P.off;
    double v[2] = {0.0, 0.0};
    const Col col = column<TY>(P, lb, rows);
    for_column<TY>(P, col, [&](const int i, const int) {
        v[0] = dmaxd(v[0], g[i]);
        v[1] += c[i] * P.dx * P.dy;
    });
    block_partial<2, 1u>(v, B.part + P.poff, nbs, lb);
}

__global__ __launch_bounds__(TPB) void k_strip_fin(Bufs B, int d, int nbs) {
    const KP P = B.kp[d];
    double v[2];
    if (!final_reduce<2, 1u>(v, B.part + P.poff, nbs)) return;
    B.S[d].gmax = v[0];
    B.S[d].removed = v[1];
}

// shift_rows row move and gamma fill (:2029-2042) into the other buffer,
// then close_mu (:789-792) on the new fields.
template <int TY>
__global__ __launch_bounds__(TPB) void k_shift(Bufs B, int d, int rows, double c_gamma) {
    const KP P = B.kp[d];
    DevScal* S = B.S + d;
    const int lb = blockIdx.x;
    const int cur = S->cur;
    const int n = P.n;
    const double* __restrict__ phi_in = B.phi[cur] + 3 * P.off;
    const double* __restrict__ c_in = B.c[cur] + P.off;
    double* __restrict__ phi_out = B.phi[cur ^ 1] + 3 * P.off;
    double* __restrict__ c_out = B.c[cur ^ 1] + P.off;
    double* __restrict__ mu_out = B.mu[cur ^ 1] + P.off;
    bool bad_mix = false, bad_c = false;
    const Col col = column<TY>(P, lb, P.ny);
    for_column<TY>(P, col, [&](const int i, const int y) {
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
    });
    if (bad_mix) atomicOr(&S->err, ERRF_MIXTURE);
    if (bad_c) atomicOr(&S->err, ERRF_NONFINITE_CARBON);
}

// ---------------------------------------------------------------------------
// Capillarity-only interface smoothing (smooth_interface_only :869-958 with
// interface_energy_only :822-841).  The force planes reuse three solver work
// planes (r, z, Ap), which are idle outside a diffusion solve.
// ---------------------------------------------------------------------------

// Obstacle + capillary force of one cell minus its mean (:896-915); the
// same expressions as the capillary part of propose_phi.
__device__ __forceinline__ void smooth_force_cell(const KP& P, int y, const double (&p)[3], const double (&qw)[3],
                                                  const double (&qe)[3], const double (&qs)[3],
                                                  const double (&qn)[3], double (&f)[3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int a1 = (a + 1) % 3, a2 = (a + 2) % 3;
        f[a] = obstacle_force(P, p, a) + P.tr * p[a1] * p[a2];
    }
#pragma unroll
    for (int dir = 0; dir < 4; ++dir) {
        if ((dir == 2 && y == 0) || (dir == 3 && y == P.ny - 1)) continue;
        const double(&q)[3] = dir == 0 ? qw : (dir == 1 ? qe : (dir == 2 ? qs : qn));
        const double dd = dir < 2 ? P.dx2 : P.dy2, rdd = dir < 2 ? P.rdx2 : P.rdy2;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = a + 1; b < 3; ++b) {
                const double wv = div_const(P.fc_p[pair_of(a, b)] * (p[a] * q[b] - q[a] * p[b]), dd, rdd);
                f[a] += wv * q[b];
                f[b] -= wv * q[a];
            }
    }
    const double mean = div_const(f[0] + f[1] + f[2], 3.0, P.r3);
#pragma unroll
    for (int a = 0; a < 3; ++a) f[a] = f[a] - mean;
}

template <int TY>
__global__ __launch_bounds__(TPB) void k_smooth_force(Bufs B, int d) {
    const KP P = B.kp[d];
    DevScal* S = B.S + d;
    const int lb = blockIdx.x;
    const int n = P.n;
    const double* __restrict__ phi = B.phi[S->cur] + 3 * P.off;
    double* __restrict__ f0 = B.r + P.off;
    double* __restrict__ f1 = B.z + P.off;
    double* __restrict__ f2 = B.Ap + P.off;
    double v[1] = {0.0};
    const Col col = column<TY>(P, lb, P.ny);
    march<TY, Phases>(
        P, col,
        [&](const int j) {
            Phases q;
            load_phi(phi, n, j, q.p);
            return q;
        },
        [&](const int i, const int y, const Phases& s, const Phases& c, const Phases& nn) {
            double qw[3], qe[3], f[3];
            load_phi(phi, n, i + col.dw, qw);
            load_phi(phi, n, i + col.de, qe);
            smooth_force_cell(P, y, c.p, qw, qe, s.p, nn.p, f);
            f0[i] = f[0];
            f1[i] = f[1];
            f2[i] = f[2];
#pragma unroll
            for (int a = 0; a < 3; ++a) v[0] = dmaxd(v[0], __builtin_fabs(f[a]));
        });
    block_partial<1, 1u>(v, B.part + P.poff, P.nb, lb);
}

__global__ __launch_bounds__(TPB) void k_smooth_force_fin(Bufs B, int d) {
    const KP P = B.kp[d];
    double v[1];
    if (final_reduce<1, 1u>(v, B.part + P.poff, P.nb)) B.S[d].sm_fmax = v[0];
}

// Trial phases of one cell: z = phi - scale*force projected onto the
// simplex (:921-931); bad when z is not finite or the projection fails.
struct Trial {
    double p[3];   // projected trial phases
    double chg;    // max_a |trial - phi|
    bool bad;
};
__device__ __forceinline__ Trial smooth_trial_at(const KP& P, const double* __restrict__ phi, const double* f0,
                                                 const double* f1, const double* f2, int n, int j, double scale) {
    Trial t;
    double p[3], z[3];
    load_phi(phi, n, j, p);
    z[0] = p[0] - scale * f0[j];
    z[1] = p[1] - scale * f1[j];
    z[2] = p[2] - scale * f2[j];
    t.bad = !(__builtin_isfinite(z[0]) && __builtin_isfinite(z[1]) && __builtin_isfinite(z[2]));
    if (!project_simplex(z, P.r3, t.p)) t.bad = true;
    t.chg = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) t.chg = dmaxd(t.chg, __builtin_fabs(t.p[a] - p[a]));
    return t;
}

// interface_energy_only of one cell (:826-838): the cell term, then the east
// and (below the top row) north face terms.
__device__ __forceinline__ double iface_cell(const KP& P, const double (&p)[3], const double (&qe)[3],
                                             const double (&qn)[3], bool north) {
    double v = obstacle_energy(P, p) + P.tr * p[0] * p[1] * p[2];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = a + 1; b < 3; ++b) {
            const double z = div_const(p[a] * qe[b] - qe[a] * p[b], P.dx, P.rdx);
            v += P.es_p[pair_of(a, b)] * z * z;
        }
    if (north) {
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = a + 1; b < 3; ++b) {
                const double z = div_const(p[a] * qn[b] - qn[a] * p[b], P.dy, P.rdy);
                v += P.es_p[pair_of(a, b)] * z * z;
            }
    }
    return v;
}

// One trial of the backtracking search (:917-937): trial phases into the
// other phi buffer, max change, failure flag and the trial's interface
// energy.  The north neighbour's trial comes from the column march, the
// east neighbour's is recomputed from phi and the force.
template <int TY>
__global__ __launch_bounds__(TPB) void k_smooth_trial(Bufs B, int d, double scale) {
    const KP P = B.kp[d];
    DevScal* S = B.S + d;
    const int lb = blockIdx.x;
    const int n = P.n;
    const int cur = S->cur;
    const double* __restrict__ phi = B.phi[cur] + 3 * P.off;
    double* __restrict__ out = B.phi[cur ^ 1] + 3 * P.off;
    const double* f0 = B.r + P.off;
    const double* f1 = B.z + P.off;
    const double* f2 = B.Ap + P.off;
    double v[3] = {0.0, 0.0, 0.0};  // max change, bad, energy
    const Col col = column<TY>(P, lb, P.ny);
    march<TY, Trial>(
        P, col, [&](const int j) { return smooth_trial_at(P, phi, f0, f1, f2, n, j, scale); },
        [&](const int i, const int y, const Trial&, const Trial& c, const Trial& nn) {
            const Trial e = smooth_trial_at(P, phi, f0, f1, f2, n, i + col.de, scale);
#pragma unroll
            for (int a = 0; a < 3; ++a) out[a * n + i] = c.p[a];
            v[0] = dmaxd(v[0], c.chg);
            if (c.bad) v[1] = 1.0;
            v[2] += iface_cell(P, c.p, e.p, nn.p, y != P.ny - 1);
        });
    block_partial<3, (1u | (1u << 2))>(v, B.part + P.poff, P.nb, lb);
}

__global__ __launch_bounds__(TPB) void k_smooth_trial_fin(Bufs B, int d) {
    const KP P = B.kp[d];
    DevScal* S = B.S + d;
    double v[3];
    if (!final_reduce<3, (1u | (1u << 2))>(v, B.part + P.poff, P.nb)) return;
    S->sm_change = v[0];
    S->sm_bad = v[1];
    S->sm_energy = v[2];
}

// interface_energy_only of the current phases.
template <int TY>
__global__ __launch_bounds__(TPB) void k_iface_energy(Bufs B, int d) {
    const KP P = B.kp[d];
    DevScal* S = B.S + d;
    const int lb = blockIdx.x;
    const int n = P.n;
    const double* __restrict__ phi = B.phi[S->cur] + 3 * P.off;
    double v[1] = {0.0};
    const Col col = column<TY>(P, lb, P.ny);
    march<TY, Phases>(
        P, col,
        [&](const int j) {
            Phases q;
            load_phi(phi, n, j, q.p);
            return q;
        },
        [&](const int i, const int y, const Phases&, const Phases& c, const Phases& nn) {
            double qe[3];
            load_phi(phi, n, i + col.de, qe);
            v[0] += iface_cell(P, c.p, qe, nn.p, y != P.ny - 1);
        });
    block_partial<1, 0u>(v, B.part + P.poff, P.nb, lb);
}

__global__ __launch_bounds__(TPB) void k_iface_energy_fin(Bufs B, int d) {
    const KP P = B.kp[d];
    double v[1];
    if (final_reduce<1, 0u>(v, B.part + P.poff, P.nb)) B.S[d].sm_energy = v[0];
}

// Clears error bits (a kernel, not a memset, so it is safe in graphs).
// Clears the ERRF_* bits `bits` of domain d (one thread), so that a call
// reports only what its own kernels flag.
__global__ void k_clear_err(Bufs B, int d, int bits) {
    if (threadIdx.x == 0) B.S[d].err &= ~bits;
}

// ---------------------------------------------------------------------------
// Initialization (initialize :1761-1837): the sharp labels of the FILL
// recipe and the carbon / chemical-potential set-up.
// ---------------------------------------------------------------------------
// One FILL operation, laid out as the C ABI's pearl_fill: the parameters in
// input order and the cosine and sine of each direction the operation uses
// (Y-junction arm k: dir[k]; pearlite arm and ellipse: dir[0]).  The host
// evaluates them with the reference's expressions, so the labels do not
// depend on the device's cos/sin.
struct FillOp {
    int method, pad;
    double p[16];
    double dir[3][2];
};
enum { FILL_CUBE = 0, FILL_CYLINDER, FILL_CENTERBOX, FILL_YJUNCTION, FILL_YJUNCTION_LAMELLAE, FILL_ADHERED_ARM,
       FILL_ELLIPSE, FILL_METHODS };

// filling_centered_lamellar_phase :1378-1389: cementite in the middle of
// every period.
__host__ __device__ inline int centred_lamella(double along, double period, double fraction, int ferrite,
                                               int cementite) {
    double pos = fmod(along, period);
    if (pos < 0.0) pos += period;
    const double width = fraction * period;
    return (pos >= 0.5 * (period - width) && pos < 0.5 * (period + width)) ? cementite : ferrite;
}

// Rotated coordinates of cell offset (ex, ey) along direction (ux, uy).
__host__ __device__ inline void along_across(double ex, double ey, double ux, double uy, double& along,
                                             double& across) {
    along = ex * ux + ey * uy;
    across = -ex * uy + ey * ux;
}

// Label of cell (gx, gy) after operation f, given its label `cur` before
// (filling_phase_at :1393-1510).  Whole-map recipes (centre box, Y
// junctions) set every cell; overlays change only the cells of their mask.
// Host-callable too, so that the labelling can be checked on the CPU.
__host__ __device__ inline int fill_label(const FillOp& f, int gx, int gy, int cur) {
    const double* p = f.p;
    constexpr double pi = 3.14159265358979323846;  // the literal of :1396
    switch (f.method) {
        case FILL_CUBE: {  // inclusive box; a gamma cube changes nothing (:1399-1405)
            const int s = static_cast<int>(p[0]);
            return (s != 2 && gx >= p[1] && gx <= p[4] && gy >= p[2] && gy <= p[5]) ? s : cur;
        }
        case FILL_CYLINDER: {  // disc of radius p[5] about (p[1], p[2]) (:1406-1410)
            const double ex = gx - p[1], ey = gy - p[2];
            const int s = static_cast<int>(p[0]);
            return (s != 2 && ex * ex + ey * ey <= p[5] * p[5]) ? s : cur;
        }
        case FILL_CENTERBOX: {  // half-open box, inside p[4], outside p[5] (:1411-1413)
            const bool in = gx >= p[0] && gx < p[1] && gy >= p[2] && gy < p[3];
            return static_cast<int>(in ? p[4] : p[5]);
        }
        case FILL_YJUNCTION:
        case FILL_YJUNCTION_LAMELLAE: {  // three arms inside a window (:1414-1438)
            const bool lam = f.method == FILL_YJUNCTION_LAMELLAE;
            const int ferrite = static_cast<int>(p[lam ? 13 : 11]);
            const int background = static_cast<int>(p[lam ? 15 : 12]);
            if (!(gx >= p[0] && gx < p[1] && gy >= p[2] && gy < p[3])) return background;
            const double ex = gx - p[4], ey = gy - p[5];
            for (int arm = 0; arm < 3; ++arm) {  // the first arm that holds the cell decides
                double along, across;
                along_across(ex, ey, f.dir[arm][0], f.dir[arm][1], along, across);
                if (along >= 0.0 && along <= p[9] && fabs(across) <= p[10]) {
                    if (!lam) return ferrite;
                    const double period = p[11] + p[12];
                    double pos = fmod(along, period);
                    if (pos < 0.0) pos += period;
                    return pos < p[11] ? ferrite : static_cast<int>(p[14]);
                }
            }
            return background;
        }
        case FILL_ADHERED_ARM: {  // lamellar arm with a sine-shaped outline (:1439-1485)
            double along, across;
            along_across(gx - p[0], gy - p[1], f.dir[0][0], f.dir[0][1], along, across);
            const double len = p[3];
            const int pairs = static_cast<int>(p[4]);
            const double period = len / static_cast<double>(pairs);
            const double fw = (1.0 - p[5]) * period, cw = p[5] * period;
            const double peak = p[6], contact = p[7], bridge = p[8];
            const int ferrite = static_cast<int>(p[9]), cementite = static_cast<int>(p[10]);
            int s = -1;
            if (along >= 0.0 && along <= len) {
                int pid = along >= len ? pairs - 1 : static_cast<int>(floor(along / period));
                pid = pid < pairs - 1 ? pid : pairs - 1;  // std::max(0, std::min(pairs - 1, pid))
                pid = pid < 0 ? 0 : pid;
                const double start = static_cast<double>(pid) * period;
                const bool first = along - start < fw;  // ferrite, then cementite within a pair
                const int seg = first ? 2 * pid : 2 * pid + 1;
                const double seg_start = first ? start : start + fw;
                s = first ? ferrite : cementite;
                double t = (along - seg_start) / (first ? fw : cw);
                if (t < 0.0) t = 0.0;
                else if (t > 1.0) t = 1.0;
                const double h0 = seg == 0 ? 0.0 : contact;
                const double h1 = seg == 2 * pairs - 1 ? 0.0 : contact;
                // sin: the device's (the reference calls std::sin); a label can
                // differ only where |across| is within an ulp of the outline.
                double half = (1.0 - t) * h0 + t * h1 + (peak - 0.5 * (h0 + h1)) * sin(pi * t);
                if (half > peak) half = peak;
                if (fabs(across) > half) s = -1;
            }
            if (bridge > 0.0) {
                const double b2 = bridge * bridge;
                if (along * along + across * across <= b2) s = ferrite;
                if ((along - len) * (along - len) + across * across <= b2) s = cementite;
            }
            return s >= 0 ? s : cur;
        }
        case FILL_ELLIPSE: {  // lamellar ellipse with optional end bridges (:1486-1508)
            double along, across;
            along_across(gx - p[0], gy - p[1], f.dir[0][0], f.dir[0][1], along, across);
            const double start = p[3], period = p[5], length = p[4] * period;
            const double semi_major = 0.5 * length, semi_minor = p[6] * semi_major;
            const double u = (along - (start + semi_major)) / semi_major, v = across / semi_minor;
            const double bridge = p[8];
            const int ferrite = static_cast<int>(p[9]), cementite = static_cast<int>(p[10]);
            int s = -1;
            if (u * u + v * v <= 1.0) s = centred_lamella(along - start, period, p[7], ferrite, cementite);
            if (bridge > 0.0) {
                const double a0 = along - start, a1 = along - (start + length);
                const double b2 = bridge * bridge;
                if (a0 * a0 + across * across <= b2 || a1 * a1 + across * across <= b2) s = ferrite;
            }
            return s >= 0 ? s : cur;
        }
        default: return cur;
    }
}

// Sharp labels of the current phases: every cell starts as gamma and the
// operations apply in order.  The operation list is read with uniform
// (scalar) loads: every lane walks the same list.
template <int TY>
__global__ __launch_bounds__(TPB) void k_fill(Bufs B, int d, const FillOp* __restrict__ ops, int count) {
    const KP P = B.kp[d];
    const int lb = blockIdx.x;
    double* __restrict__ phi = B.phi[B.S[d].cur] + 3 * P.off;
    const int n = P.n;
    const Col col = column<TY>(P, lb, P.ny);
    for_column<TY>(P, col, [&](const int i, const int y) {
        int s = 2;
        for (int k = 0; k < count; ++k) s = fill_label(ops[k], col.x, y, s);
        phi[i] = s == 0 ? 1.0 : 0.0;
        phi[n + i] = s == 1 ? 1.0 : 0.0;
        phi[2 * n + i] = s == 2 ? 1.0 : 0.0;
    });
}

// Diffuse-profile phases of the current state (initialize's lamella, flat
// and circle branches, :1767-1771, :1787-1791) from host profile values:
// kind 0 col[x], row[y]; kind 1 row[y]; kind 2 cell[i].
template <int TY>
__global__ __launch_bounds__(TPB) void k_profile(Bufs B, int d, int kind, const double* __restrict__ col,
                                                 const double* __restrict__ row, const double* __restrict__ cell,
                                                 int pa, int pb) {
    const KP P = B.kp[d];
    const int lb = blockIdx.x;
    double* __restrict__ phi = B.phi[B.S[d].cur] + 3 * P.off;
    const int n = P.n;
    const Col c = column<TY>(P, lb, P.ny);
    for_column<TY>(P, c, [&](const int i, const int y) {
        double p[3] = {0.0, 0.0, 0.0};
        if (kind == 0) {
            const double th = col[c.x], product = row[y];
            p[0] = (1 - th) * product;
            p[1] = th * product;
            p[2] = 1 - product;
        } else {
            const double u = kind == 1 ? row[y] : cell[i];
            p[pa] = u;
            p[pb] = 1 - u;
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) phi[a * n + i] = p[a];
    });
}

// Sums over the cells of the carbon intercept and the susceptibility of
// the current phases (the mean_carbon branch of initialize :1814-1822),
// blocked order; q2 flags a cell whose mixture the reference rejects.
template <int TY>
__global__ __launch_bounds__(TPB) void k_init_sums(Bufs B, int d) {
    const KP P = B.kp[d];
    const int lb = blockIdx.x;
    const double* __restrict__ phi = B.phi[B.S[d].cur] + 3 * P.off;
    double v[3] = {0.0, 0.0, 0.0};
    const Col col = column<TY>(P, lb, P.ny);
    for_column<TY>(P, col, [&](const int i, const int) {
        double p[3];
        load_phi(phi, P.n, i, p);
        Mix m;
        if (!mixture(P, p, m)) v[2] = 1.0;
        v[0] += m.c0;
        v[1] += m.s;
    });
    block_partial<3, (1u << 4)>(v, B.part + P.poff, P.nb, lb);
}

__global__ __launch_bounds__(TPB) void k_init_sums_fin(Bufs B, int d) {
    const KP P = B.kp[d];
    DevScal* S = B.S + d;
    double v[3];
    if (!final_reduce<3, (1u << 4)>(v, B.part + P.poff, P.nb)) return;
    S->in_c0 = v[0];
    S->in_s = v[1];
    S->in_bad = v[2];
}

// Uniform chemical potential mu and the carbon of the current phases at it
// (initialize :1823-1826): c = composition(mu, mixture(phi)).
template <int TY>
__global__ __launch_bounds__(TPB) void k_init_carbon(Bufs B, int d, double mu) {
    const KP P = B.kp[d];
    const int lb = blockIdx.x;
    const int cur = B.S[d].cur;
    const double* __restrict__ phi = B.phi[cur] + 3 * P.off;
    double* __restrict__ c = B.c[cur] + P.off;
    double* __restrict__ m_out = B.mu[cur] + P.off;
    const Col col = column<TY>(P, lb, P.ny);
    for_column<TY>(P, col, [&](const int i, const int) {
        double p[3];
        load_phi(phi, P.n, i, p);
        Mix m;
        mixture(P, p, m);  // an invalid mixture is reported by the state check that follows
        c[i] = composition(mu, m);
        m_out[i] = mu;
    });
}

}  // namespace pearl_gpu
