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
// kernels -- begin, state reductions, propose, setup, residual, PCG slots,
// finalize, decide -- and every kernel acts only on the domains whose phase
// and PCG state call for it.  A round therefore has no host decision in it:
// it can be replayed from a captured graph, and one launch serves any number
// of independent domains of different sizes.
#pragma once

#include <hip/hip_runtime.h>

namespace pearl_gpu {

constexpr int TPB = 256;        // threads per block of every cell kernel
constexpr int MAX_SLOTS = 64;   // per-launch work records (profiling)
constexpr int NKIND = 16;       // per-domain run counters

// Per-domain parameter block.  Geometry of the domain inside the packed
// arena, then the model constants, all derived on the host exactly as the
// reference forms them.
struct KP {
    int nx, ny, n;          // grid; n = nx*ny cells
    int nb;                 // blocks of this domain in a cell-kernel grid
    int boff;               // first block of this domain in the grid
    long long off;          // first cell of this domain in every plane
    long long poff;         // start of this domain's reduction partials
    double dx, dy, dx2, dy2;  // dx2 = dx*dx (the reference's d.dx*d.dx)
    double A[3], B[3], C[3], D[3];
    double chi[3];          // 1.0/(2.0*A)        Parabola::chi   :84
    double qpsi[3];         // -0.25/A             Parabola::psi   :87
    double twoA[3];         // 2.0*A               Parabola::c_at_mu :83
    double obs, tr;         // 16*sigma/(pi*pi*eps), triple/eps   :964
    double face_coef;       // 2*eps*sigma                        :981
    double eps_sigma;       // eps*sigma                          :816
    double eps, sigma, Vm;
    double tau[3], tau_min;
    double psiref[3];       // psi_a(mu_ref) - psi_gamma(mu_ref), psiref[2]=0  :797
    double mu_ref;
    // Controls (:767-773)
    double dt_max, dt_min, max_phase_change, cg_rtol, cg_atol, closure_tol, energy_rtol, energy_atol;
    double mass_floor;      // gx*gy*dx*dy*1e-6                         :1152
    double energy_scale;    // sigma*max(gx*dx, gy*dy)                  :1158
    int cg_max, max_retries, check_energy, phase_only;
};

// Controller phases of one domain.
enum { PH_IDLE = 0, PH_START = 1, PH_PROPOSE = 2, PH_SOLVE = 3, PH_DECIDE = 4 };
// PCG states (implicit_diffusion :1047-1080).  CG_ITER: take a step;
// CG_TRUE: recompute the true residual (:1053); CG_OK / CG_FAIL: done;
// CG_IDLE: not started.
enum { CG_IDLE = 0, CG_ITER = 1, CG_TRUE = 2, CG_OK = 3, CG_FAIL = 4 };
// Device error codes (the C ABI's PEARL_ERR_* values).
enum { ST_OK = 0, ST_INVALID_TIME = 1, ST_DT_MIN = 2, ST_RETRY_LIMIT = 3, ST_MIXTURE = 4, ST_PROJECTION = 5 };
enum { ERRF_MIXTURE = 1, ERRF_PROJECTION = 2, ERRF_NONFINITE_CARBON = 4, ERRF_MIXTURE_STATE = 8 };
// Kernels counted per domain (work actually done).
enum { RUN_PROPOSE = 0, RUN_SETUP, RUN_RESIDUAL0, RUN_CG_APPLY, RUN_CG_UPDATE, RUN_TRUE_RESIDUAL, RUN_FINALIZE,
       RUN_STATE, RUN_COUNT };

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
    double dt;             // dt of the current attempt
    double e0, m0;
    // attempt
    int live;              // the phase increment passed (:1131)
    int cg_state, cg_it, restart, pidx;
    int err;               // ERRF_* bits
    unsigned counter;      // last-block counter
    double change;
    double rhs2, rz, rr, pAp, alpha, beta, tol, cg_res;
    double closure, valid, mass, energy;
    double e_valid, e_mass, e_energy;
    double front, gmax, removed;
    double dmin_c, dmax_c, dmin_mu, dmax_mu, dsumerr, dmass, denergy;
    // StepInfo of the last accepted step
    double i_dt, i_change, i_residual, i_closure, i_energy_change;
    int i_cg, i_retries;
    // counters
    long long attempts, rejected_attempts, cg_all, cg_total;
    int cg_peak, pad;
    unsigned runs[NKIND];
};

__device__ __forceinline__ double dmaxd(double a, double b) { return (a < b) ? b : a; }  // std::max
__device__ __forceinline__ double dmind(double a, double b) { return (b < a) ? b : a; }  // std::min

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
template <int NQ, unsigned OPS>
__device__ __forceinline__ void block_partial(double (&v)[NQ], double* __restrict__ part, int nb, int lb) {
    __shared__ double sh[4][NQ];
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
            part[q * nb + lb] = comb(op, comb(op, sh[0][q], sh[1][q]), comb(op, sh[2][q], sh[3][q]));
        }
    }
}

// Stage 3 in a finisher kernel (one 256-thread block per domain, launched
// after the cell kernel, so its partials are visible without fences):
// thread t folds partials t, t+256, ... then the same tree.  Returns true in
// thread 0, with the totals in v.
template <int NQ, unsigned OPS>
__device__ __forceinline__ bool final_reduce(double (&v)[NQ], const double* __restrict__ part, int nb) {
    __shared__ double sh[4][NQ];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int op = op_of<OPS>(q);
        double acc = ident(op);
        for (int b = threadIdx.x; b < nb; b += TPB) acc = comb(op, acc, part[q * nb + b]);
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
// Packed arena and domain lookup
// ---------------------------------------------------------------------------
struct Bufs {
    double* phi[2];   // domain d: phi[b] + 3*off + a*n + i
    double* c[2];
    double* mu[2];
    double *chi, *icpt, *K, *rhs, *x, *diag, *r, *z, *p0, *p1, *Ap;
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
__device__ __forceinline__ Where where(const Bufs& B, int only_d) {
    if (only_d >= 0) return {only_d, static_cast<int>(blockIdx.x)};
    const int d = B.blk2dom[blockIdx.x];
    return {d, static_cast<int>(blockIdx.x) - B.kp[d].boff};
}

__device__ __forceinline__ void record(const Bufs& B, int slot, int lb, int kind) {
    if (slot >= 0 && lb == 0 && threadIdx.x == 0) B.kinds[slot] = kind;
}

#define PEARL_CELL_LOOP(KK, P, lb)                                            \
    _Pragma("unroll") for (int j_ = 0; j_ < (KK); ++j_) {                     \
        const int i = ((lb) * (KK) + j_) * TPB + static_cast<int>(threadIdx.x); \
        if (i >= (P).n) break;

#define PEARL_CELL_LOOP_END }

// Loop head of implicit_diffusion (:1047-1067) for iteration `it` with the
// current r.r: info.it, info.residual, then the convergence and
// iteration-limit branches.  One thread.
__device__ __forceinline__ void cg_head(const KP& P, DevScal* S, int it) {
    const double res = __builtin_sqrt(S->rr);
    S->cg_it = it;
    S->cg_res = res;
    if (!__builtin_isfinite(res)) S->cg_state = CG_FAIL;
    else if (res <= S->tol) S->cg_state = CG_TRUE;
    else if (it == P.cg_max) S->cg_state = CG_FAIL;
    else S->cg_state = CG_ITER;
    if (S->cg_state == CG_FAIL) S->phase = PH_DECIDE;
}

// ---------------------------------------------------------------------------
// Round kernels
// ---------------------------------------------------------------------------

// Start of a step (:1120-1123) and reset of the per-attempt flags.  One
// thread per domain.
__global__ void k_begin(Bufs B) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= B.ndom) return;
    const KP& P = B.kp[d];
    DevScal* S = B.S + d;
    if (S->phase == PH_START) {
        double dt = S->dt_state;
        if (P.dt_max < dt) dt = P.dt_max;
        if (S->stop_time - S->time < dt) dt = S->stop_time - S->time;
        if (!(dt > 0) || !__builtin_isfinite(dt)) {
            S->status = ST_INVALID_TIME;
            S->phase = PH_IDLE;
            return;
        }
        S->dt = dt;
        S->attempt = 0;
        S->phase = PH_PROPOSE;
        S->need_state = ((P.check_energy && !S->e0_valid) || (!P.phase_only && !S->m0_valid)) ? 1 : 0;
    }
    if (S->phase == PH_PROPOSE) {
        S->live = 0;
        S->cg_state = CG_IDLE;
        S->cg_it = 0;
        S->cg_res = 0;
        S->closure = 0;
        S->restart = 1;
        S->pidx = 0;
        S->pad = 0;  // "stepped": k_cg_a took a step that k_cg_b must complete
    }
}

// valid_state / total_carbon / reduced_energy of one cell of a state
// (:1096-1116, :784-788, :793-821).
__device__ __forceinline__ void state_cell(const KP& P, const double* phi, const double* c, const double* mu, int i,
                                           bool& good, double& mass, double& energy, bool& mix_ok) {
    const Nbr nb = neighbours(P, i);
    double e;
    const double ci = c[i];
    accept_cell(P, phi, i, nb, ci, mu[i], good, e, mix_ok);
    mass += ci;
    energy += e;
}

// e0 and mass0 of a step (:1122-1123) when they are not cached.
template <int KK>
__global__ __launch_bounds__(TPB) void k_state(Bufs B, int slot) {
    const Where w = where(B, -1);
    const KP& P = B.kp[w.d];
    DevScal* S = B.S + w.d;
    if (!(S->phase == PH_PROPOSE && S->need_state)) return;
    record(B, slot, w.lb, 1);
    const int cur = S->cur;
    const double* phi = B.phi[cur] + 3 * P.off;
    const double* c = B.c[cur] + P.off;
    const double* mu = B.mu[cur] + P.off;
    double v[4] = {1.0, 0.0, 0.0, 0.0};
    bool good = true, mix_ok = true;
    PEARL_CELL_LOOP(KK, P, w.lb)
    state_cell(P, phi, c, mu, i, good, v[1], v[2], mix_ok);
    PEARL_CELL_LOOP_END
    v[0] = good ? 1.0 : 0.0;
    v[3] = mix_ok ? 0.0 : 1.0;
    block_partial<4, (2u | (1u << 6))>(v, B.part + P.poff, P.nb, w.lb);
}

__global__ __launch_bounds__(TPB) void k_state_fin(Bufs B) {
    const int d = blockIdx.x;
    const KP& P = B.kp[d];
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

// P1: propose_phi :961-1002, then the phase-increment test (:1131).
template <int KK>
__global__ __launch_bounds__(TPB) void k_propose(Bufs B, int slot) {
    const Where w = where(B, -1);
    const KP& P = B.kp[w.d];
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
    PEARL_CELL_LOOP(KK, P, w.lb)
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
                const double wv = P.face_coef * (p[a] * q[b] - q[a] * p[b]) / dd;
                force[a] += wv * q[b];
                force[b] -= wv * q[a];
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
    if (!project_simplex(z, zp)) v[1] = 1.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) v[0] = dmaxd(v[0], __builtin_fabs(zp[a] - p[a]));
#pragma unroll
    for (int a = 0; a < 3; ++a) out[a * n + i] = zp[a];
    PEARL_CELL_LOOP_END
    block_partial<2, (1u | (1u << 2))>(v, B.part + P.poff, P.nb, w.lb);
}

__global__ __launch_bounds__(TPB) void k_propose_fin(Bufs B) {
    const int d = blockIdx.x;
    const KP& P = B.kp[d];
    DevScal* S = B.S + d;
    if (S->phase != PH_PROPOSE || S->need_state) return;
    double v[2];
    if (!final_reduce<2, (1u | (1u << 2))>(v, B.part + P.poff, P.nb)) return;
    S->runs[RUN_PROPOSE]++;
    S->change = v[0];
    if (v[1] > 0) {
        S->status = ST_PROJECTION;
        S->phase = PH_IDLE;
        return;
    }
    // :1131 -- the rest of the attempt runs only if the increment passes
    S->live = (__builtin_isfinite(v[0]) && v[0] <= P.max_phase_change) ? 1 : 0;
    S->phase = S->live ? PH_SOLVE : PH_DECIDE;
}

// P2: implicit_diffusion setup :1013-1020.
template <int KK>
__global__ __launch_bounds__(TPB) void k_setup(Bufs B, int slot) {
    const Where w = where(B, -1);
    const KP& P = B.kp[w.d];
    DevScal* S = B.S + w.d;
    if (!(S->phase == PH_SOLVE && S->cg_state == CG_IDLE) || P.phase_only) return;
    record(B, slot, w.lb, 1);
    if (w.lb == 0 && threadIdx.x == 0) S->runs[RUN_SETUP]++;
    const int cur = S->cur;
    const double* __restrict__ phi_t = B.phi[cur ^ 1] + 3 * P.off;
    const double* __restrict__ c_old = B.c[cur] + P.off;
    const double* __restrict__ mu_old = B.mu[cur] + P.off;
    double* __restrict__ chi = B.chi + P.off;
    double* __restrict__ icpt = B.icpt + P.off;
    double* __restrict__ K = B.K + P.off;
    double* __restrict__ rhs = B.rhs + P.off;
    double* __restrict__ x = B.x + P.off;
    bool bad = false;
    PEARL_CELL_LOOP(KK, P, w.lb)
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
__global__ __launch_bounds__(TPB) void k_residual0(Bufs B, int slot) {
    const Where w = where(B, -1);
    const KP& P = B.kp[w.d];
    DevScal* S = B.S + w.d;
    if (!(S->phase == PH_SOLVE && S->cg_state == CG_IDLE) || P.phase_only) return;
    record(B, slot, w.lb, 1);
    const double dt = S->dt;
    const double* __restrict__ chi = B.chi + P.off;
    const double* __restrict__ K = B.K + P.off;
    const double* __restrict__ rhs = B.rhs + P.off;
    const double* __restrict__ x = B.x + P.off;
    double* __restrict__ diag = B.diag + P.off;
    double* __restrict__ r = B.r + P.off;
    double* __restrict__ z = B.z + P.off;
    double v[3] = {0.0, 0.0, 0.0};
    PEARL_CELL_LOOP(KK, P, w.lb)
    const Nbr nb = neighbours(P, i);
    double wf[4];
    face_weights(P, K, i, nb, wf);
    const double c = chi[i];
    const double ax = apply_cell(P, x, c, dt, i, nb, wf);
    const double d = c + dt * (wf[0] + wf[1] + wf[2] + wf[3]);
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
    block_partial<3, 0u>(v, B.part + P.poff, P.nb, w.lb);
}

__global__ __launch_bounds__(TPB) void k_residual0_fin(Bufs B) {
    const int d = blockIdx.x;
    const KP& P = B.kp[d];
    DevScal* S = B.S + d;
    if (!(S->phase == PH_SOLVE && S->cg_state == CG_IDLE) || P.phase_only) return;
    double v[3];
    if (!final_reduce<3, 0u>(v, B.part + P.poff, P.nb)) return;
    S->runs[RUN_RESIDUAL0]++;
    if (S->err & ERRF_MIXTURE) {
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
    cg_head(P, S, 0);
}

// PCG kernel A.  In state CG_ITER: the search direction of this iteration
// p = z + beta p_old (:1079; p = z after a (re)start :1042/:1062), kept in
// the other p buffer, then Ap = A p (:1068) and p.Ap; the finisher checks
// for breakdown (:1070) and forms the step length (:1071).  In state
// CG_TRUE: the true residual r = rhs - A x with z = r/diag (:1053-1062);
// the finisher accepts (:1056) or restarts from it (:1060-1065).
template <int KK>
__global__ __launch_bounds__(TPB) void k_cg_a(Bufs B, int slot) {
    const Where w = where(B, -1);
    const KP& P = B.kp[w.d];
    DevScal* S = B.S + w.d;
    if (S->phase != PH_SOLVE) return;
    const int state = S->cg_state;
    const double dt = S->dt;
    const double* __restrict__ chi = B.chi + P.off;
    const double* __restrict__ K = B.K + P.off;
    double* __restrict__ z = B.z + P.off;
    if (state == CG_ITER) {
        record(B, slot, w.lb, 1);
        const int pi = S->pidx;
        const double* __restrict__ po = (pi ? B.p1 : B.p0) + P.off;
        double* __restrict__ pn = (pi ? B.p0 : B.p1) + P.off;
        double* __restrict__ Ap = B.Ap + P.off;
        const bool fresh = S->restart != 0;
        const double beta = S->beta;
        double v[1] = {0.0};
        PEARL_CELL_LOOP(KK, P, w.lb)
        const Nbr nb = neighbours(P, i);
        double wf[4];
        face_weights(P, K, i, nb, wf);
        const double pc = fresh ? z[i] : z[i] + beta * po[i];
        const double pw = fresh ? z[nb.w] : z[nb.w] + beta * po[nb.w];
        const double pe = fresh ? z[nb.e] : z[nb.e] + beta * po[nb.e];
        const double ps = fresh ? z[nb.s] : z[nb.s] + beta * po[nb.s];
        const double pq = fresh ? z[nb.nn] : z[nb.nn] + beta * po[nb.nn];
        const double a = chi[i] * pc + dt * (wf[0] * (pc - pw) + wf[1] * (pc - pe) + wf[2] * (pc - ps) + wf[3] * (pc - pq));
        pn[i] = pc;
        Ap[i] = a;
        v[0] += pc * a;
        PEARL_CELL_LOOP_END
        block_partial<1, 0u>(v, B.part + P.poff, P.nb, w.lb);
    } else if (state == CG_TRUE) {
        record(B, slot, w.lb, 2);
        const double* __restrict__ rhs = B.rhs + P.off;
        const double* __restrict__ x = B.x + P.off;
        const double* __restrict__ diag = B.diag + P.off;
        double* __restrict__ r = B.r + P.off;
        double v[2] = {0.0, 0.0};
        PEARL_CELL_LOOP(KK, P, w.lb)
        const Nbr nb = neighbours(P, i);
        double wf[4];
        face_weights(P, K, i, nb, wf);
        const double rr = rhs[i] - apply_cell(P, x, chi[i], dt, i, nb, wf);
        const double zz = rr / diag[i];
        r[i] = rr;
        z[i] = zz;
        v[0] += rr * rr;
        v[1] += rr * zz;
        PEARL_CELL_LOOP_END
        block_partial<2, 0u>(v, B.part + P.poff, P.nb, w.lb);
    }
}

__global__ __launch_bounds__(TPB) void k_cg_a_fin(Bufs B) {
    const int d = blockIdx.x;
    const KP& P = B.kp[d];
    DevScal* S = B.S + d;
    if (S->phase != PH_SOLVE) return;
    const int state = S->cg_state;
    if (state == CG_ITER) {
        double v[1];
        if (!final_reduce<1, 0u>(v, B.part + P.poff, P.nb)) return;
        S->runs[RUN_CG_APPLY]++;
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
        } else {
            S->rz = v[1];
            S->restart = 1;
            cg_head(P, S, S->cg_it + 1);
        }
    }
}

// PCG kernel B, after a step of kernel A: x += alpha p; r -= alpha Ap;
// z = r/diag; sums r.z and r.r; the finisher forms beta (:1077) and
// evaluates the loop head of the next iteration.
template <int KK>
__global__ __launch_bounds__(TPB) void k_cg_b(Bufs B, int slot) {
    const Where w = where(B, -1);
    const KP& P = B.kp[w.d];
    DevScal* S = B.S + w.d;
    if (!(S->phase == PH_SOLVE && S->cg_state == CG_ITER && S->pad == 1)) return;
    record(B, slot, w.lb, 1);
    const double* __restrict__ p = (S->pidx ? B.p1 : B.p0) + P.off;
    double* __restrict__ x = B.x + P.off;
    double* __restrict__ r = B.r + P.off;
    double* __restrict__ z = B.z + P.off;
    const double* __restrict__ Ap = B.Ap + P.off;
    const double* __restrict__ diag = B.diag + P.off;
    const double alpha = S->alpha;
    double v[2] = {0.0, 0.0};
    PEARL_CELL_LOOP(KK, P, w.lb)
    x[i] += alpha * p[i];
    const double rr = r[i] - alpha * Ap[i];
    const double zz = rr / diag[i];
    r[i] = rr;
    z[i] = zz;
    v[0] += rr * zz;
    v[1] += rr * rr;
    PEARL_CELL_LOOP_END
    block_partial<2, 0u>(v, B.part + P.poff, P.nb, w.lb);
}

__global__ __launch_bounds__(TPB) void k_cg_b_fin(Bufs B) {
    const int d = blockIdx.x;
    const KP& P = B.kp[d];
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
    cg_head(P, S, S->cg_it + 1);
}

// F + A: flux-form carbon update, closure, mu reconstruction (:1082-1093),
// fused with the acceptance sweep on the trial state: valid_state (:1096),
// total_carbon (:784) and reduced_energy (:793).
template <int KK>
__global__ __launch_bounds__(TPB) void k_finalize(Bufs B, int slot) {
    const Where w = where(B, -1);
    const KP& P = B.kp[w.d];
    DevScal* S = B.S + w.d;
    if (!(S->phase == PH_SOLVE && S->cg_state == CG_OK)) return;
    record(B, slot, w.lb, 1);
    const int cur = S->cur;
    const double dt = S->dt;
    const double* __restrict__ phi_t = B.phi[cur ^ 1] + 3 * P.off;
    const double* __restrict__ chi = B.chi + P.off;
    const double* __restrict__ icpt = B.icpt + P.off;
    const double* __restrict__ K = B.K + P.off;
    const double* __restrict__ x = B.x + P.off;
    const double* __restrict__ c_old = B.c[cur] + P.off;
    double* __restrict__ c_new = B.c[cur ^ 1] + P.off;
    double* __restrict__ mu_new = B.mu[cur ^ 1] + P.off;
    double v[4] = {0.0, 1.0, 0.0, 0.0};
    bool good = true, mix_ok = true;
    PEARL_CELL_LOOP(KK, P, w.lb)
    const Nbr nb = neighbours(P, i);
    double wf[4];
    face_weights(P, K, i, nb, wf);
    const double xi = x[i];
    const double change = dt * (wf[0] * (x[nb.w] - xi) + wf[1] * (x[nb.e] - xi) + wf[2] * (x[nb.s] - xi) + wf[3] * (x[nb.nn] - xi));
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
    block_partial<4, (1u | (2u << 2))>(v, B.part + P.poff, P.nb, w.lb);
}

__global__ __launch_bounds__(TPB) void k_finalize_fin(Bufs B) {
    const int d = blockIdx.x;
    const KP& P = B.kp[d];
    DevScal* S = B.S + d;
    if (!(S->phase == PH_SOLVE && S->cg_state == CG_OK)) return;
    double v[4];
    if (!final_reduce<4, (1u | (2u << 2))>(v, B.part + P.poff, P.nb)) return;
    S->runs[RUN_FINALIZE]++;
    S->closure = v[0];
    S->valid = v[1];
    S->mass = v[2];
    S->energy = v[3];
    S->phase = PH_DECIDE;
}

// Phase-only acceptance (:1134-1136): carbon = composition(mu, mixture(phi')),
// mu unchanged; then valid_state and reduced_energy on the trial state.
template <int KK>
__global__ __launch_bounds__(TPB) void k_phase_only(Bufs B, int slot) {
    const Where w = where(B, -1);
    const KP& P = B.kp[w.d];
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
    PEARL_CELL_LOOP(KK, P, w.lb)
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
    v[4] = mix_ok ? 0.0 : 1.0;
    block_partial<5, (1u | (2u << 2) | (1u << 8))>(v, B.part + P.poff, P.nb, w.lb);
}

__global__ __launch_bounds__(TPB) void k_phase_only_fin(Bufs B) {
    const int d = blockIdx.x;
    const KP& P = B.kp[d];
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

// Acceptance tests and the dt controller of advance (:1131-1173).  One
// thread per domain.
__global__ void k_decide(Bufs B) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= B.ndom) return;
    const KP& P = B.kp[d];
    DevScal* S = B.S + d;
    if (S->phase != PH_DECIDE) return;
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
        S->phase = (S->taken >= S->target_steps || !(S->time < S->stop_time)) ? PH_IDLE : PH_START;
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

__global__ void k_clear_kinds(int* kinds, int n) {
    const int i = threadIdx.x;
    if (i < n) kinds[i] = 0;
}

// ---------------------------------------------------------------------------
// Direct (single-domain) kernels: state reductions, diagnostics, window.
// ---------------------------------------------------------------------------

// valid_state, total_carbon and reduced_energy of domain d's current state.
template <int KK>
__global__ __launch_bounds__(TPB) void k_state_direct(Bufs B, int d) {
    const KP& P = B.kp[d];
    DevScal* S = B.S + d;
    const int lb = blockIdx.x;
    const int cur = S->cur;
    const double* phi = B.phi[cur] + 3 * P.off;
    const double* c = B.c[cur] + P.off;
    const double* mu = B.mu[cur] + P.off;
    double v[4] = {1.0, 0.0, 0.0, 0.0};
    bool good = true, mix_ok = true;
    PEARL_CELL_LOOP(KK, P, lb)
    state_cell(P, phi, c, mu, i, good, v[1], v[2], mix_ok);
    PEARL_CELL_LOOP_END
    v[0] = good ? 1.0 : 0.0;
    v[3] = mix_ok ? 0.0 : 1.0;
    block_partial<4, (2u | (1u << 6))>(v, B.part + P.poff, P.nb, lb);
}

__global__ __launch_bounds__(TPB) void k_state_direct_fin(Bufs B, int d) {
    const KP& P = B.kp[d];
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
template <int KK>
__global__ __launch_bounds__(TPB) void k_diagnostics(Bufs B, int d) {
    const KP& P = B.kp[d];
    DevScal* S = B.S + d;
    const int lb = blockIdx.x;
    const int cur = S->cur;
    const double* phi = B.phi[cur] + 3 * P.off;
    const double* c = B.c[cur] + P.off;
    const double* mu = B.mu[cur] + P.off;
    double v[8] = {1e300, -1e300, 1e300, -1e300, 0.0, 0.0, 0.0, 0.0};
    bool good = true, mix_ok = true;
    PEARL_CELL_LOOP(KK, P, lb)
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
    v[7] = mix_ok ? 0.0 : 1.0;
    block_partial<8, DIAG_OPS>(v, B.part + P.poff, P.nb, lb);
}

__global__ __launch_bounds__(TPB) void k_diagnostics_fin(Bufs B, int d) {
    const KP& P = B.kp[d];
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
template <int KK>
__global__ __launch_bounds__(TPB) void k_front(Bufs B, int d) {
    const KP& P = B.kp[d];
    DevScal* S = B.S + d;
    const int lb = blockIdx.x;
    const double* g = B.phi[S->cur] + 3 * P.off + 2 * static_cast<long long>(P.n);
    double v[1] = {-1.0};
    PEARL_CELL_LOOP(KK, P, lb)
    const int y = i / P.nx;
    if (y < P.ny - 1) {
        const double a = 1 - 2 * g[i], b = 1 - 2 * g[i + P.nx];
        if (a >= 0 && b < 0) v[0] = dmaxd(v[0], ((y + 1) - .5 + a / (a - b)) * P.dy);
    }
    PEARL_CELL_LOOP_END
    block_partial<1, 1u>(v, B.part + P.poff, P.nb, lb);
}

__global__ __launch_bounds__(TPB) void k_front_fin(Bufs B, int d) {
    const KP& P = B.kp[d];
    double v[1];
    if (final_reduce<1, 1u>(v, B.part + P.poff, P.nb)) B.S[d].front = v[0];
}

// Max gamma and removed carbon (c*dx)*dy over the first m cells (:2028,
// :2067).  Launched with nbs = ceil(m/(256*KK)) blocks.
template <int KK>
__global__ __launch_bounds__(TPB) void k_strip(Bufs B, int d, int m, int nbs) {
    const KP& P = B.kp[d];
    DevScal* S = B.S + d;
    const int cur = S->cur;
    const double* g = B.phi[cur] + 3 * P.off + 2 * static_cast<long long>(P.n);
    const double* c = B.c[cur] + P.off;
    double v[2] = {0.0, 0.0};
    _Pragma("unroll") for (int j_ = 0; j_ < KK; ++j_) {
        const int i = (static_cast<int>(blockIdx.x) * KK + j_) * TPB + static_cast<int>(threadIdx.x);
        if (i >= m) break;
        v[0] = dmaxd(v[0], g[i]);
        v[1] += c[i] * P.dx * P.dy;
    }
    block_partial<2, 1u>(v, B.part + P.poff, nbs, blockIdx.x);
}

__global__ __launch_bounds__(TPB) void k_strip_fin(Bufs B, int d, int nbs) {
    const KP& P = B.kp[d];
    double v[2];
    if (!final_reduce<2, 1u>(v, B.part + P.poff, nbs)) return;
    B.S[d].gmax = v[0];
    B.S[d].removed = v[1];
}

// shift_rows row move and gamma fill (:2029-2042) into the other buffer,
// then close_mu (:789-792) on the new fields.
template <int KK>
__global__ __launch_bounds__(TPB) void k_shift(Bufs B, int d, int rows, double c_gamma) {
    const KP& P = B.kp[d];
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
    PEARL_CELL_LOOP(KK, P, lb)
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

// Clears error bits (a kernel, not a memset, so it is safe in graphs).
__global__ void k_clear_err(Bufs B, int d) { B.S[d].err = 0; }

}  // namespace pearl_gpu
