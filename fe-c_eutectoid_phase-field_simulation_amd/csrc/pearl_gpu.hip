// C ABI of the FP64 per-step solver (include/pearl_gpu.h): context,
// device-resident state, and the host-side step controller that restates
// advance (pearl_serial.cpp:1119-1174), implicit_diffusion's PCG control
// (:1044-1081) and maybe_shift (:2053-2071) on top of the kernels in
// pearl_kernels.hpp.  The controller reads back only the few scalars each
// decision needs; all field work stays on the device.
#include "pearl_gpu.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "pearl_kernels.hpp"

using namespace pearl_gpu;

namespace {

const char* kWhy[5] = {"phase increment", "linear/closure", "composition/phase bounds", "mass conservation",
                       "free-energy increase"};

const char* kKernelNames[PEARL_KERNEL_COUNT] = {
    "propose_phi", "diffusion_setup", "cg_residual0", "cg_apply", "cg_update", "true_residual",
    "idle", "finalize_accept", "state_reduce", "moving_window"};

// Algorithmic FP64 words per cell of each kernel (each logical array counted
// once per pass): propose R phi(3),mu W phi'(3); setup R phi'(3),c,mu
// W chi,icpt,K,rhs,x; residual0 R x,K,chi,rhs W r,diag,z; apply R z,p_old,K,
// chi W p,Ap; update R x,p,r,Ap,diag W x,r,z; true residual R x,K,chi,rhs,diag
// W r,z; finalize+accept R phi'(3),x,K,c,icpt,chi W c',mu'; state reduce
// R phi(3),c,mu.
const double kWords[PEARL_KERNEL_COUNT] = {7, 10, 7, 6, 8, 7, 0, 10, 5, 0};

// Launch accounting: a launch whose work depends on a device-side decision
// is resolved after the next synchronisation.
enum LaunchCond { COND_ALWAYS = 0, COND_LIVE = 1, COND_CG_OK = 2, COND_SLOT = 3 };

}  // namespace

struct pearl_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    pearl_params prm{};
    double mu_ref = 0;
    KP kp{};
    int n = 0, k = 1;
    // device buffers
    double* mem = nullptr;
    double *phi[2]{}, *c[2]{}, *mu[2]{};
    int cur = 0;
    double *chi = nullptr, *icpt = nullptr, *K = nullptr, *rhs = nullptr, *x = nullptr, *diag = nullptr;
    double *r = nullptr, *z = nullptr, *p0 = nullptr, *p1 = nullptr, *Ap = nullptr, *part = nullptr;
    int cg_slots = 6;  // PCG slots enqueued per synchronisation
    DevScal* S = nullptr;
    DevScal* hS = nullptr;  // pinned host copy
    pearl_scalars sc{};
    bool e0_valid = false, m0_valid = false;
    double e0 = 0, m0 = 0;
    std::string err;
    long long attempts = 0, rejected_attempts = 0, cg_iterations = 0;
    // profiling
    bool prof = false;
    struct Pending {
        int kid, cond, slot;
        double cells;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> free_events;
    long long launches[PEARL_KERNEL_COUNT]{};
    double ms[PEARL_KERNEL_COUNT]{};
    double bytes[PEARL_KERNEL_COUNT]{};
};

namespace {

int fail(pearl_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

int hip_fail(pearl_ctx* c, hipError_t e, const char* where) {
    return fail(c, PEARL_ERR_DEVICE, std::string("HIP error in ") + where + ": " + hipGetErrorString(e));
}

#define HIP_CHECK(ctx, call)                                 \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #call); \
    } while (0)

hipEvent_t get_event(pearl_ctx* c) {
    if (!c->free_events.empty()) {
        hipEvent_t e = c->free_events.back();
        c->free_events.pop_back();
        return e;
    }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
}

// Resolve the launches since the last synchronisation (the stream is
// idle): a launch counts under its kernel, with its algorithmic bytes, if it
// did work, and under PEARL_KERNEL_CG_IDLE otherwise.
void harvest(pearl_ctx* c) {
    const DevScal* h = c->hS;
    for (auto& pd : c->pending) {
        int kid = pd.kid;
        bool real = true;
        if (pd.cond == COND_LIVE) real = h->live != 0;
        else if (pd.cond == COND_CG_OK) real = pd.slot == 1;
        else if (pd.cond == COND_SLOT) {
            const int kind = h->kind[pd.slot];
            real = kind != KIND_NOOP;
            if (kid == PEARL_KERNEL_CG_APPLY && kind == KIND_TRUE) kid = PEARL_KERNEL_TRUE_RESIDUAL;
        }
        if (!real) kid = PEARL_KERNEL_CG_IDLE;
        c->launches[kid]++;
        c->bytes[kid] += kWords[kid] * 8.0 * pd.cells;
        if (pd.a) {
            float t = 0;
            (void)hipEventElapsedTime(&t, pd.a, pd.b);
            c->ms[kid] += t;
            c->free_events.push_back(pd.a);
            c->free_events.push_back(pd.b);
        }
    }
    c->pending.clear();
}

// Brackets one launch: records it for accounting and, when profiling, with
// a pair of events on the context's stream.
struct ProfScope {
    pearl_ctx* c;
    size_t idx;
    ProfScope(pearl_ctx* cc, int kid, double cells, int cond = COND_ALWAYS, int slot = 0) : c(cc) {
        pearl_ctx::Pending pd{kid, cond, slot, cells, nullptr, nullptr};
        if (c->prof) {
            pd.a = get_event(c);
            pd.b = get_event(c);
            (void)hipEventRecord(pd.a, c->stream);
        }
        c->pending.push_back(pd);
        idx = c->pending.size() - 1;
    }
    ~ProfScope() {
        if (c->prof) (void)hipEventRecord(c->pending[idx].b, c->stream);
    }
};

dim3 grid_of(const pearl_ctx* c) { return dim3(static_cast<unsigned>(c->kp.nb)); }

#define PEARL_DISPATCH(ctx, kname, grid, ...)                                                          \
    do {                                                                                              \
        switch ((ctx)->k) {                                                                           \
            case 1: hipLaunchKernelGGL(kname<1>, grid, dim3(TPB), 0, (ctx)->stream, __VA_ARGS__); break; \
            case 2: hipLaunchKernelGGL(kname<2>, grid, dim3(TPB), 0, (ctx)->stream, __VA_ARGS__); break; \
            case 4: hipLaunchKernelGGL(kname<4>, grid, dim3(TPB), 0, (ctx)->stream, __VA_ARGS__); break; \
            default: hipLaunchKernelGGL(kname<8>, grid, dim3(TPB), 0, (ctx)->stream, __VA_ARGS__); break; \
        }                                                                                             \
    } while (0)

// Copy the device scalars to the pinned host block and wait.
int sync_scalars(pearl_ctx* c) {
    HIP_CHECK(c, hipMemcpyAsync(c->hS, c->S, sizeof(DevScal), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    HIP_CHECK(c, hipGetLastError());
    harvest(c);
    return PEARL_OK;
}

int err_from_flags(pearl_ctx* c, int flags) {
    if (flags & ERRF_PROJECTION) return fail(c, PEARL_ERR_PROJECTION, "Simplex projection failed");
    if (flags & ERRF_MIXTURE) return fail(c, PEARL_ERR_MIXTURE, "Invalid mixture coefficients");
    if (flags & ERRF_NONFINITE_CARBON) return fail(c, PEARL_ERR_NONFINITE_CARBON, "Nonfinite total carbon");
    return PEARL_OK;
}

double* cur_phi(pearl_ctx* c) { return c->phi[c->cur]; }
double* cur_c(pearl_ctx* c) { return c->c[c->cur]; }
double* cur_mu(pearl_ctx* c) { return c->mu[c->cur]; }
double* trial_phi(pearl_ctx* c) { return c->phi[1 - c->cur]; }
double* trial_c(pearl_ctx* c) { return c->c[1 - c->cur]; }
double* trial_mu(pearl_ctx* c) { return c->mu[1 - c->cur]; }

// valid_state / total_carbon / reduced_energy of the current state.
int state_reduce(pearl_ctx* c, bool* valid, double* mass, double* energy, bool mixture_is_error) {
    {
        ProfScope ps(c, PEARL_KERNEL_STATE_REDUCE, c->n);
        PEARL_DISPATCH(c, k_state_reduce, grid_of(c), c->kp, cur_phi(c), cur_c(c), cur_mu(c), c->part, c->S);
    }
    int st = sync_scalars(c);
    if (st) return st;
    if (mixture_is_error && (c->hS->err & ERRF_MIXTURE_STATE)) return err_from_flags(c, ERRF_MIXTURE);
    if (c->hS->err) {
        c->hS->err = 0;
        hipLaunchKernelGGL(k_clear_flags, dim3(1), dim3(1), 0, c->stream, c->S);
    }
    if (valid) *valid = c->hS->e_valid > 0.5;
    if (mass) *mass = c->hS->e_mass * c->kp.dx * c->kp.dy;
    if (energy) *energy = c->hS->e_energy * c->kp.dx * c->kp.dy;
    return PEARL_OK;
}

void launch_propose(pearl_ctx* c, double dt) {
    ProfScope ps(c, PEARL_KERNEL_PROPOSE, c->n);
    PEARL_DISPATCH(c, k_propose, grid_of(c), c->kp, cur_phi(c), cur_mu(c), trial_phi(c), dt, c->part, c->S);
}

int propose(pearl_ctx* c, double dt, double* change) {
    launch_propose(c, dt);
    int st = sync_scalars(c);
    if (st) return st;
    if (c->hS->err & ERRF_PROJECTION) return err_from_flags(c, ERRF_PROJECTION);
    *change = c->hS->change;
    return PEARL_OK;
}

// implicit_diffusion :1010-1094 on (cur c, cur mu, trial phi) -> trial c, mu,
// fused with the acceptance reductions of the trial state.  The PCG loop is
// controlled on the device (k_residual0 / k_cg_a / k_cg_b finishers); the
// host enqueues `cg_slots` iterations at a time, each predicated on the
// device state, and synchronises once per batch.
struct DiffOut {
    int ok = 0, it = 0;
    double residual = 0, closure = 0;
    bool finalized = false;
};

void launch_setup(pearl_ctx* c, double dt) {
    const KP& P = c->kp;
    const dim3 g = grid_of(c);
    {
        ProfScope ps(c, PEARL_KERNEL_SETUP, c->n, COND_LIVE);
        PEARL_DISPATCH(c, k_setup, g, P, trial_phi(c), cur_c(c), cur_mu(c), c->chi, c->icpt, c->K, c->rhs, c->x, c->S);
    }
    {
        ProfScope ps(c, PEARL_KERNEL_RESIDUAL0, c->n, COND_LIVE);
        PEARL_DISPATCH(c, k_residual0, g, P, c->chi, c->K, c->rhs, c->x, c->diag, c->r, c->z, dt, c->part, c->S);
    }
}

// Enqueue `slots` PCG iterations and the (predicated) finalize.
void launch_cg_batch(pearl_ctx* c, double dt, int slots) {
    const KP& P = c->kp;
    const dim3 g = grid_of(c);
    for (int j = 0; j < slots; ++j) {
        {
            ProfScope ps(c, PEARL_KERNEL_CG_APPLY, c->n, COND_SLOT, 2 * j);
            PEARL_DISPATCH(c, k_cg_a, g, P, c->chi, c->K, c->rhs, c->x, c->diag, c->r, c->z, c->p0, c->p1, c->Ap, dt,
                           2 * j, c->part, c->S);
        }
        {
            ProfScope ps(c, PEARL_KERNEL_CG_UPDATE, c->n, COND_SLOT, 2 * j + 1);
            PEARL_DISPATCH(c, k_cg_b, g, P, c->x, c->r, c->z, c->p0, c->p1, c->Ap, c->diag, 2 * j + 1, c->part, c->S);
        }
    }
    ProfScope ps(c, PEARL_KERNEL_FINALIZE, c->n, COND_CG_OK, 0);
    PEARL_DISPATCH(c, k_finalize, g, P, trial_phi(c), c->chi, c->icpt, c->K, c->x, cur_c(c), trial_c(c), trial_mu(c),
                   dt, c->part, c->S);
}

// After the first synchronisation of an attempt: finish the PCG solve if
// the enqueued iterations were not enough, and read its outcome.
int sync_attempt(pearl_ctx* c);

int finish_diffusion(pearl_ctx* c, double dt, DiffOut& out) {
    while (c->hS->cg_state == CG_ITER || c->hS->cg_state == CG_TRUE) {
        launch_cg_batch(c, dt, c->cg_slots);
        const int st = sync_attempt(c);
        if (st) return st;
    }
    const DevScal* h = c->hS;
    out.it = h->cg_it;
    out.residual = h->cg_res;
    c->cg_iterations += h->cg_it;
    c->cg_slots = std::min(28, std::max(2, h->cg_it + 2));
    if (h->cg_state != CG_OK) return PEARL_OK;
    out.closure = h->closure;
    out.ok = out.closure <= c->prm.closure_tol;
    out.finalized = true;
    return PEARL_OK;
}

// Marks the last finalize launch of the pending list as the one that ran.
void mark_final_finalize(pearl_ctx* c) {
    for (auto it = c->pending.rbegin(); it != c->pending.rend(); ++it)
        if (it->cond == COND_CG_OK) {
            it->slot = 1;
            break;
        }
}

// Synchronise after a batch.  The finalize launch runs only in the batch in
// which the solve finished, and it is the last one enqueued in that batch.
int sync_attempt(pearl_ctx* c) {
    HIP_CHECK(c, hipMemcpyAsync(c->hS, c->S, sizeof(DevScal), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    HIP_CHECK(c, hipGetLastError());
    if (c->hS->cg_state == CG_OK) mark_final_finalize(c);
    harvest(c);
    return PEARL_OK;
}

int diffusion(pearl_ctx* c, double dt, DiffOut& out) {
    hipLaunchKernelGGL(k_begin_solve, dim3(1), dim3(1), 0, c->stream, c->S);
    launch_setup(c, dt);
    launch_cg_batch(c, dt, c->cg_slots);
    int st = sync_attempt(c);
    if (st) return st;
    if (c->hS->err & ERRF_MIXTURE) return err_from_flags(c, ERRF_MIXTURE);
    return finish_diffusion(c, dt, out);
}

int advance_impl(pearl_ctx* c, double stop_time, pearl_step_info* info) {
    const pearl_params& prm = c->prm;
    pearl_scalars& s = c->sc;
    double dt = s.dt;
    if (prm.dt_max < dt) dt = prm.dt_max;
    if (stop_time - s.time < dt) dt = stop_time - s.time;
    std::memset(info, 0, sizeof(*info));
    info->why = -1;
    if (!(dt > 0) || !std::isfinite(dt)) return fail(c, PEARL_ERR_INVALID_TIME, "invalid advance time");
    bool need_state = (prm.check_energy && !c->e0_valid) || (!prm.phase_only && !c->m0_valid);
    int why = PEARL_WHY_PHASE;
    for (int attempt = 0; attempt <= prm.max_retries; ++attempt) {
        c->attempts++;
        pearl_step_info out{};
        out.retries = attempt;
        out.accepted_dt = dt;
        out.why = -1;
        // One batch: [state reductions] propose, setup, residual0, PCG
        // iterations, finalize; everything after propose is predicated on
        // the device-side phase-increment decision.
        if (need_state) {
            ProfScope ps(c, PEARL_KERNEL_STATE_REDUCE, c->n);
            PEARL_DISPATCH(c, k_state_reduce, grid_of(c), c->kp, cur_phi(c), cur_c(c), cur_mu(c), c->part, c->S);
        }
        launch_propose(c, dt);
        if (prm.phase_only) {
            ProfScope ps(c, PEARL_KERNEL_FINALIZE, c->n, COND_LIVE);
            PEARL_DISPATCH(c, k_phase_only, grid_of(c), c->kp, trial_phi(c), cur_mu(c), trial_c(c), trial_mu(c),
                           c->part, c->S);
        } else {
            launch_setup(c, dt);
            launch_cg_batch(c, dt, c->cg_slots);
        }
        int st = sync_attempt(c);
        if (st) return st;
        const DevScal* h = c->hS;
        if (need_state) {
            if (prm.check_energy && (h->err & ERRF_MIXTURE_STATE)) return err_from_flags(c, ERRF_MIXTURE);
            if (prm.check_energy) {
                c->e0 = h->e_energy * c->kp.dx * c->kp.dy;
                c->e0_valid = true;
            }
            c->m0 = h->e_mass * c->kp.dx * c->kp.dy;
            c->m0_valid = true;
            need_state = false;
        }
        if (h->err & ERRF_PROJECTION) return err_from_flags(c, ERRF_PROJECTION);
        const double e0 = prm.check_energy ? c->e0 : 0;
        const double mass0 = prm.phase_only ? 0 : c->m0;
        const double change = h->change;
        out.max_phase_change = change;
        bool good = std::isfinite(change) && change <= prm.max_phase_change;
        why = PEARL_WHY_PHASE;
        double mass = 0, energy = 0, valid = 0;
        if (good) {
            if (h->err & ERRF_MIXTURE) return err_from_flags(c, ERRF_MIXTURE);
            if (!prm.phase_only) {
                DiffOut d;
                st = finish_diffusion(c, dt, d);
                if (st) return st;
                good = d.ok != 0;
                out.cg_iterations = d.it;
                out.linear_residual = d.residual;
                out.closure = d.closure;
                why = PEARL_WHY_LINEAR;
            }
            // (sum*dx)*dy as total_carbon :787 and reduced_energy :820
            valid = c->hS->valid;
            mass = c->hS->mass * c->kp.dx * c->kp.dy;
            energy = c->hS->energy * c->kp.dx * c->kp.dy;
        }
        if (good) {
            good = valid > 0.5;
            why = PEARL_WHY_BOUNDS;
        }
        if (good && !prm.phase_only) {
            const double lim = prm.nx * prm.ny * prm.dx * prm.dy * 1e-6;
            good = std::fabs(mass - mass0) <= 1e-11 * std::max(std::fabs(mass0), lim);
            why = PEARL_WHY_MASS;
        }
        if (good && prm.check_energy) {
            const double e1 = energy;
            out.energy_change = e1 - e0;
            const double tol = prm.energy_atol + prm.energy_rtol * std::max(std::fabs(e0), prm.sigma * std::max(prm.nx * prm.dx, prm.ny * prm.dy));
            good = std::isfinite(e1) && e1 <= e0 + tol;
            why = PEARL_WHY_ENERGY;
        }
        if (good) {
            c->cur = 1 - c->cur;
            s.time = s.time + dt;
            s.step = s.step + 1;
            s.rejected = s.rejected + attempt;
            s.dt = std::min(prm.dt_max, dt * 1.1);
            if (prm.check_energy) {
                c->e0 = energy;
                c->e0_valid = true;
            }
            if (!prm.phase_only) {
                c->m0 = mass;
                c->m0_valid = true;
            }
            *info = out;
            return PEARL_OK;
        }
        c->rejected_attempts++;
        dt *= .5;
        if (dt < prm.dt_min) {
            info->why = why;
            return fail(c, PEARL_ERR_DT_MIN, std::string("step rejected below dt_min: ") + kWhy[why]);
        }
    }
    info->why = why;
    return fail(c, PEARL_ERR_RETRY_LIMIT, std::string("step retry limit: ") + kWhy[why]);
}

bool params_ok(const pearl_params* p) {
    if (!p) return false;
    if (p->nx < 3 || p->ny < 2 || !(p->dx > 0) || !(p->dy > 0)) return false;
    if (static_cast<long long>(p->nx) * p->ny >= (1LL << 31) - (1LL << 20)) return false;
    for (int i = 0; i < 3; ++i)
        if (!(p->A[i] > 0) || !std::isfinite(p->A[i]) || !(p->tau[i] > 0)) return false;
    if (!(p->eps > 0) || !(p->sigma > 0) || !(p->Vm > 0) || p->triple < 0) return false;
    return true;
}

}  // namespace

extern "C" {

int pearl_create(const pearl_params* params, double mu_reference, int precision, int device, int cells_per_thread,
                 pearl_ctx** out) {
    if (!out) return PEARL_ERR_ARGUMENT;
    *out = nullptr;
    if (!params_ok(params)) return PEARL_ERR_ARGUMENT;
    if (precision != PEARL_FP64) return PEARL_ERR_ARGUMENT;
    auto* c = new pearl_ctx();
    c->device = device;
    c->prm = *params;
    c->mu_ref = mu_reference;
    c->n = params->nx * params->ny;
    int k = cells_per_thread;
    if (k <= 0) k = c->n >= (1 << 21) ? 4 : 1;
    if (k != 1 && k != 2 && k != 4 && k != 8) {
        delete c;
        return PEARL_ERR_ARGUMENT;
    }
    c->k = k;
    if (hipSetDevice(device) != hipSuccess) {
        delete c;
        return PEARL_ERR_DEVICE;
    }
    // Kernel parameter block: every derived constant uses the reference's
    // own expression.
    KP& P = c->kp;
    P.nx = params->nx;
    P.ny = params->ny;
    P.n = c->n;
    P.nb = (c->n + TPB * k - 1) / (TPB * k);
    P.dx = params->dx;
    P.dy = params->dy;
    P.dx2 = params->dx * params->dx;
    P.dy2 = params->dy * params->dy;
    for (int i = 0; i < 3; ++i) {
        P.A[i] = params->A[i];
        P.B[i] = params->B[i];
        P.C[i] = params->C[i];
        P.D[i] = params->D[i];
        P.chi[i] = 1.0 / (2.0 * params->A[i]);
        P.qpsi[i] = -0.25 / params->A[i];
        P.twoA[i] = 2.0 * params->A[i];
        P.tau[i] = params->tau[i];
    }
    P.tau_min = params->tau[0];
    if (params->tau[1] < P.tau_min) P.tau_min = params->tau[1];
    if (params->tau[2] < P.tau_min) P.tau_min = params->tau[2];
    const double eps = params->eps, sigma = params->sigma;
    P.obs = 16 * sigma / (std::acos(-1.) * std::acos(-1.) * eps);
    P.tr = params->triple / eps;
    P.face_coef = 2 * eps * sigma;
    P.eps_sigma = eps * sigma;
    P.eps = eps;
    P.Vm = params->Vm;
    auto psi = [&](int p, double mu) {
        const double d = mu - params->B[p];
        return std::fma(-0.25 / params->A[p], d * d, params->C[p]);
    };
    P.psiref[0] = psi(0, mu_reference) - psi(2, mu_reference);
    P.psiref[1] = psi(1, mu_reference) - psi(2, mu_reference);
    P.psiref[2] = 0;
    P.mu_ref = mu_reference;
    P.max_phase_change = params->max_phase_change;
    P.cg_rtol = params->cg_rtol;
    P.cg_atol = params->cg_atol;
    P.cg_max = params->cg_max;
    P.phase_only = params->phase_only;

    const size_t n = static_cast<size_t>(c->n);
    const size_t plane = (n + 31) & ~static_cast<size_t>(31);  // 256-byte aligned planes
    const size_t nparts = 8 * static_cast<size_t>(P.nb) + 64;
    const size_t total = plane * 21 + nparts;
    if (hipMalloc(&c->mem, total * sizeof(double)) != hipSuccess) {
        delete c;
        return PEARL_ERR_ALLOC;
    }
    double* q = c->mem;
    for (int b = 0; b < 2; ++b) {
        c->phi[b] = q;
        q += 3 * plane;
    }
    for (int b = 0; b < 2; ++b) {
        c->c[b] = q;
        q += plane;
    }
    for (int b = 0; b < 2; ++b) {
        c->mu[b] = q;
        q += plane;
    }
    c->chi = q; q += plane;
    c->icpt = q; q += plane;
    c->K = q; q += plane;
    c->rhs = q; q += plane;
    c->x = q; q += plane;
    c->diag = q; q += plane;
    c->r = q; q += plane;
    c->z = q; q += plane;
    c->p0 = q; q += plane;
    c->p1 = q; q += plane;
    c->Ap = q; q += plane;
    c->part = q;
    if (hipMalloc(&c->S, sizeof(DevScal)) != hipSuccess || hipHostMalloc(&c->hS, sizeof(DevScal)) != hipSuccess) {
        pearl_destroy(c);
        return PEARL_ERR_ALLOC;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        pearl_destroy(c);
        return PEARL_ERR_DEVICE;
    }
    (void)hipMemsetAsync(c->S, 0, sizeof(DevScal), c->stream);
    (void)hipMemsetAsync(c->mem, 0, total * sizeof(double), c->stream);
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        pearl_destroy(c);
        return PEARL_ERR_DEVICE;
    }
    std::memset(c->hS, 0, sizeof(DevScal));
    *out = c;
    return PEARL_OK;
}

int pearl_destroy(pearl_ctx* c) {
    if (!c) return PEARL_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& pd : c->pending) {
        (void)hipEventDestroy(pd.a);
        (void)hipEventDestroy(pd.b);
    }
    for (auto e : c->free_events) (void)hipEventDestroy(e);
    if (c->mem) (void)hipFree(c->mem);
    if (c->S) (void)hipFree(c->S);
    if (c->hS) (void)hipHostFree(c->hS);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return PEARL_OK;
}

const char* pearl_last_error(const pearl_ctx* c) { return c ? c->err.c_str() : "null context"; }

int pearl_set_state(pearl_ctx* c, const double* phi, const double* carbon, const double* mu,
                    const pearl_scalars* scalars) {
    if (!c) return PEARL_ERR_ARGUMENT;
    (void)hipSetDevice(c->device);
    const size_t n = static_cast<size_t>(c->n);
    if (phi) HIP_CHECK(c, hipMemcpyAsync(cur_phi(c), phi, 3 * n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (carbon) HIP_CHECK(c, hipMemcpyAsync(cur_c(c), carbon, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (mu) HIP_CHECK(c, hipMemcpyAsync(cur_mu(c), mu, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    if (scalars) c->sc = *scalars;
    c->e0_valid = false;
    c->m0_valid = false;
    return PEARL_OK;
}

int pearl_get_state(pearl_ctx* c, double* phi, double* carbon, double* mu, pearl_scalars* scalars) {
    if (!c) return PEARL_ERR_ARGUMENT;
    (void)hipSetDevice(c->device);
    const size_t n = static_cast<size_t>(c->n);
    if (phi) HIP_CHECK(c, hipMemcpyAsync(phi, cur_phi(c), 3 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (carbon) HIP_CHECK(c, hipMemcpyAsync(carbon, cur_c(c), n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (mu) HIP_CHECK(c, hipMemcpyAsync(mu, cur_mu(c), n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    if (scalars) *scalars = c->sc;
    return PEARL_OK;
}

int pearl_advance(pearl_ctx* c, double stop_time, pearl_step_info* info) {
    if (!c || !info) return PEARL_ERR_ARGUMENT;
    (void)hipSetDevice(c->device);
    return advance_impl(c, stop_time, info);
}

int pearl_run(pearl_ctx* c, double stop_time, long long max_steps, long long* taken, long long* cg_total,
              int* cg_peak, pearl_step_info* last) {
    if (!c) return PEARL_ERR_ARGUMENT;
    (void)hipSetDevice(c->device);
    long long t = 0;
    pearl_step_info info{};
    int st = PEARL_OK;
    while (c->sc.time < stop_time && t < max_steps) {
        st = advance_impl(c, stop_time, &info);
        if (st) break;
        ++t;
        if (cg_total) *cg_total += info.cg_iterations;
        if (cg_peak) *cg_peak = std::max(*cg_peak, info.cg_iterations);
    }
    if (taken) *taken = t;
    if (last) *last = info;
    return st;
}

int pearl_maybe_shift(pearl_ctx* c, int moving_window, double trigger, double target, double c_gamma, int* shifted,
                      double* front) {
    if (!c || !shifted) return PEARL_ERR_ARGUMENT;
    (void)hipSetDevice(c->device);
    *shifted = 0;
    if (front) *front = -1;
    if (!moving_window) return PEARL_OK;
    const KP& P = c->kp;
    {
        ProfScope ps(c, PEARL_KERNEL_WINDOW, c->n);
        PEARL_DISPATCH(c, k_front, grid_of(c), P, cur_phi(c), c->part, c->S);
    }
    int st = sync_scalars(c);
    if (st) return st;
    const double fr = c->hS->front;
    if (front) *front = fr;
    if (fr < trigger) return PEARL_OK;
    const int rows = static_cast<int>(std::floor((fr - target) / P.dy));
    const int scan_rows = std::max(0, std::min(rows, P.ny));
    double removed = 0;
    if (scan_rows > 0) {
        const int m = scan_rows * P.nx;
        const int nbs = (m + TPB * c->k - 1) / (TPB * c->k);
        {
            ProfScope ps(c, PEARL_KERNEL_WINDOW, m);
            PEARL_DISPATCH(c, k_strip, dim3(nbs), P, cur_phi(c), cur_c(c), m, nbs, c->part, c->S);
        }
        st = sync_scalars(c);
        if (st) return st;
        if (c->hS->gmax > 1e-6) return fail(c, PEARL_ERR_UNSAFE_SHIFT, "unsafe shift: removed strip contains gamma");
        removed = c->hS->removed;
    }
    if (rows <= 0 || rows >= P.ny || c_gamma < 0 || c_gamma > 1) return fail(c, PEARL_ERR_INVALID_SHIFT, "invalid shift");
    const double added = rows * P.nx * c_gamma * P.dx * P.dy;
    {
        ProfScope ps(c, PEARL_KERNEL_WINDOW, c->n);
        PEARL_DISPATCH(c, k_shift, grid_of(c), P, cur_phi(c), cur_c(c), trial_phi(c), trial_c(c), trial_mu(c), rows,
                       c_gamma, c->S);
    }
    st = sync_scalars(c);
    if (st) return st;
    if (c->hS->err) {
        const int f = c->hS->err;
        return err_from_flags(c, f);
    }
    c->cur = 1 - c->cur;
    c->sc.shift_distance += rows * P.dy;
    c->sc.shifts++;
    c->sc.carbon_added += added;
    c->sc.carbon_removed += removed;
    c->e0_valid = false;
    c->m0_valid = false;
    *shifted = 1;
    return PEARL_OK;
}

int pearl_diagnostics(pearl_ctx* c, pearl_diag* out) {
    if (!c || !out) return PEARL_ERR_ARGUMENT;
    (void)hipSetDevice(c->device);
    {
        ProfScope ps(c, PEARL_KERNEL_STATE_REDUCE, c->n);
        PEARL_DISPATCH(c, k_diagnostics, grid_of(c), c->kp, cur_phi(c), cur_c(c), cur_mu(c), c->part, c->S);
    }
    int st = sync_scalars(c);
    if (st) return st;
    if (c->hS->err & ERRF_MIXTURE) return err_from_flags(c, ERRF_MIXTURE);
    out->c_min = c->hS->dmin_c;
    out->c_max = c->hS->dmax_c;
    out->mu_min = c->hS->dmin_mu;
    out->mu_max = c->hS->dmax_mu;
    out->phi_sum_error = c->hS->dsumerr;
    out->mass = c->hS->dmass * c->kp.dx * c->kp.dy;
    out->energy = c->hS->denergy * c->kp.dx * c->kp.dy;
    return PEARL_OK;
}

int pearl_propose_phi(pearl_ctx* c, double dt, double* phi_out, double* max_change) {
    if (!c || !max_change) return PEARL_ERR_ARGUMENT;
    (void)hipSetDevice(c->device);
    int st = propose(c, dt, max_change);
    if (st) return st;
    if (phi_out) {
        HIP_CHECK(c, hipMemcpyAsync(phi_out, trial_phi(c), 3 * static_cast<size_t>(c->n) * sizeof(double),
                                    hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(c, hipStreamSynchronize(c->stream));
    }
    return PEARL_OK;
}

int pearl_implicit_diffusion(pearl_ctx* c, const double* phi_trial, double dt, double* carbon_out, double* mu_out,
                             pearl_diffusion_info* info) {
    if (!c || !phi_trial || !info) return PEARL_ERR_ARGUMENT;
    (void)hipSetDevice(c->device);
    const size_t n = static_cast<size_t>(c->n);
    HIP_CHECK(c, hipMemcpyAsync(trial_phi(c), phi_trial, 3 * n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    DiffOut d;
    int st = diffusion(c, dt, d);
    if (st) return st;
    info->ok = d.ok;
    info->it = d.it;
    info->residual = d.residual;
    info->closure = d.closure;
    const double* srcc = d.finalized ? trial_c(c) : cur_c(c);
    const double* srcm = d.finalized ? trial_mu(c) : cur_mu(c);
    if (carbon_out) HIP_CHECK(c, hipMemcpyAsync(carbon_out, srcc, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (mu_out) HIP_CHECK(c, hipMemcpyAsync(mu_out, srcm, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return PEARL_OK;
}

int pearl_reduced_energy(pearl_ctx* c, double* energy) {
    if (!c || !energy) return PEARL_ERR_ARGUMENT;
    (void)hipSetDevice(c->device);
    return state_reduce(c, nullptr, nullptr, energy, !c->prm.phase_only);
}

int pearl_total_carbon(pearl_ctx* c, double* mass) {
    if (!c || !mass) return PEARL_ERR_ARGUMENT;
    (void)hipSetDevice(c->device);
    return state_reduce(c, nullptr, mass, nullptr, false);
}

int pearl_valid_state(pearl_ctx* c, int* valid) {
    if (!c || !valid) return PEARL_ERR_ARGUMENT;
    (void)hipSetDevice(c->device);
    bool v = false;
    int st = state_reduce(c, &v, nullptr, nullptr, false);
    *valid = v ? 1 : 0;
    return st;
}

int pearl_get_layout(const pearl_ctx* c, int* tpb, int* cpt) {
    if (!c) return PEARL_ERR_ARGUMENT;
    if (tpb) *tpb = TPB;
    if (cpt) *cpt = c->k;
    return PEARL_OK;
}

int pearl_profile(pearl_ctx* c, int enable) {
    if (!c) return PEARL_ERR_ARGUMENT;
    (void)hipSetDevice(c->device);
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    harvest(c);
    c->prof = enable != 0;
    for (int i = 0; i < PEARL_KERNEL_COUNT; ++i) {
        c->launches[i] = 0;
        c->ms[i] = 0;
        c->bytes[i] = 0;
    }
    return PEARL_OK;
}

int pearl_kernel_stats(pearl_ctx* c, int kernel, long long* launches, double* total_ms, double* bytes) {
    if (!c || kernel < 0 || kernel >= PEARL_KERNEL_COUNT) return PEARL_ERR_ARGUMENT;
    (void)hipSetDevice(c->device);
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    harvest(c);
    if (launches) *launches = c->launches[kernel];
    if (total_ms) *total_ms = c->ms[kernel];
    if (bytes) *bytes = c->bytes[kernel];
    return PEARL_OK;
}

const char* pearl_kernel_name(int kernel) {
    if (kernel < 0 || kernel >= PEARL_KERNEL_COUNT) return "";
    return kKernelNames[kernel];
}

int pearl_counters(const pearl_ctx* c, long long* attempts, long long* rejected, long long* cg) {
    if (!c) return PEARL_ERR_ARGUMENT;
    if (attempts) *attempts = c->attempts;
    if (rejected) *rejected = c->rejected_attempts;
    if (cg) *cg = c->cg_iterations;
    return PEARL_OK;
}

}  // extern "C"
