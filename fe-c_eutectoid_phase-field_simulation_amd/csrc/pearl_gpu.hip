// C ABI of the per-step solver (include/pearl_gpu.h).
//
// A context owns one or more domains packed in one HBM arena.  advance()
// (pearl_serial.cpp:1119-1174) runs on the device: the host only enqueues
// "rounds" of predicated kernels (pearl_kernels.hpp) and reads the device
// scalars when a call has to return.  A round is captured once into a HIP
// graph and replayed; with profiling on, rounds are launched eagerly and
// bracketed by events so that each launch can be attributed to a kernel.
#include "pearl_gpu.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <new>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "pearl_adjoint.hpp"
#include "pearl_kernels.hpp"

using namespace pearl_gpu;

namespace {

const char* kWhy[5] = {"phase increment", "linear/closure", "composition/phase bounds", "mass conservation",
                       "free-energy increase"};

const char* kKernelNames[PEARL_KERNEL_COUNT] = {
    "propose_phi", "diffusion_prepare", "cg_apply", "cg_update", "true_residual",
    "idle", "finalize_accept", "state_reduce", "moving_window", "reduce_finish", "cg_apply_first"};

// Algorithmic bytes per cell of the work each device run counter counts
// (each logical array once per pass), the kernel id it is reported under,
// and whether it counts launches (RUN_FIN_LAG adds the bytes of a finalize
// launch already counted).  Plain form: 8 B per value; mixed form: diag, z
// and p are 4 B.
//   propose             R phi(3),mu             W phi'(3)
//   prepare             R phi'(3),c,mu          W chi,K,diag,r,p (= z)
//   apply (with x step) R z,p_old,K,chi,x       W p,Ap,x     (x: mu before its first store)
//   apply (fresh)       R p,K,chi               W Ap
//   update              R r,Ap,diag             W r,z        (the adjoint's also R x,p W x)
//   true residual       R x,K,chi,rhs,diag      W r,z        (the adjoint's; the forward's forms rhs)
//   finalize+accept     R phi'(3),x,K,c         W c',mu'     (+ R p when it applies the last step)
//   state reduce        R phi(3),c,mu
struct RunBytes {
    int kid;
    double bytes[2];
    bool launch;
};
const RunBytes kRunBytes[RUN_COUNT] = {
    {PEARL_KERNEL_PROPOSE, {56, 56}, true},       {PEARL_KERNEL_PREPARE, {80, 72}, true},
    {PEARL_KERNEL_CG_APPLY, {64, 52}, true},      {PEARL_KERNEL_CG_UPDATE, {40, 32}, true},
    {PEARL_KERNEL_TRUE_RESIDUAL, {56, 48}, true}, {PEARL_KERNEL_FINALIZE, {64, 64}, true},
    {PEARL_KERNEL_STATE_REDUCE, {40, 40}, true},  {PEARL_KERNEL_CG_FIRST, {32, 28}, true},
    {PEARL_KERNEL_FINALIZE, {8, 4}, false},
};

}  // namespace

struct pearl_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int ndom = 0, k = 1;
    int mx = 0;  // precision form: PEARL_FP64, PEARL_MIXED or PEARL_FP32_STENCIL (both FP32-storage forms are "mixed")
    std::vector<pearl_params> prm;
    std::vector<KP> kp;             // host copy of the parameter blocks
    std::vector<pearl_scalars> sc;  // host-side copy of the State scalars
    double stop_eps = 0;            // pearl_set_stop_tolerance
    double* d_dt1 = nullptr;        // device copy of pearl_step_fixed's dt
    double* d_dtn = nullptr;        // device copy of pearl_run_fixed's dts
    long long d_dtn_cap = 0;
    long long total_cells = 0, total_blocks = 0, total_parts = 0;
    // device memory
    double* mem = nullptr;
    Bufs B{};
    KP* d_kp = nullptr;
    DevScal* d_S = nullptr;
    int* d_blk2dom = nullptr;
    int* d_kinds = nullptr;
    DevScal* hS = nullptr;  // pinned mirror
    std::string err;
    int cg_slots = 6;  // PCG slots per round
    // graphs of one round: [0] plain, [1] with the state-reduction kernels,
    // [2] with the fixed-run state copy (adjoint segment recompute)
    // captured rounds: [0] plain, [1] plain after a leading begin, [2] with
    // the state reductions after a leading begin, [3] ending in the fixed-run
    // state copy, [4] the same after a leading begin
    hipGraph_t graph[5] = {};
    hipGraphExec_t graph_exec[5] = {};
    int graph_slots[5] = {};
    bool any_phase_only = false;  // a phase-only domain: rounds carry its kernels
    // profiling
    bool prof = false;
    long long launches[PEARL_KERNEL_COUNT]{};
    double ms[PEARL_KERNEL_COUNT]{};
    double bytes[PEARL_KERNEL_COUNT]{};
    std::vector<unsigned> runs_seen;  // per domain x NKIND at the last harvest
    struct Ev {
        int kid, slot;
        hipEvent_t a, b;
    };
    std::vector<Ev> pending;
    std::vector<hipEvent_t> free_events;
};

namespace {

int fail(pearl_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

// Host allocation failed inside an entry point: no exception crosses the C ABI.
int alloc_fail(pearl_ctx* c) noexcept {
    if (c) {
        try {
            c->err = "host allocation failed";
        } catch (...) {
        }
    }
    return PEARL_ERR_ALLOC;
}

int hip_fail(pearl_ctx* c, hipError_t e, const char* where) {
    // Reset the runtime's last error, so that a later call's launch check
    // does not report this one again (a sticky device fault stays sticky).
    (void)hipGetLastError();
    return fail(c, PEARL_ERR_DEVICE, std::string("HIP error in ") + where + ": " + hipGetErrorString(e));
}

#define HIP_CHECK(ctx, call)                                   \
    do {                                                       \
        hipError_t e_ = (call);                                \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #call); \
    } while (0)

#define PEARL_LAUNCH(ctx, kname, grid, ...)                                                              \
    do {                                                                                                 \
        switch ((ctx)->k) {                                                                              \
            case 1: hipLaunchKernelGGL(kname<1>, grid, dim3(TPB), 0, (ctx)->stream, __VA_ARGS__); break;  \
            case 2: hipLaunchKernelGGL(kname<2>, grid, dim3(TPB), 0, (ctx)->stream, __VA_ARGS__); break;  \
            case 4: hipLaunchKernelGGL(kname<4>, grid, dim3(TPB), 0, (ctx)->stream, __VA_ARGS__); break;  \
            case 16: hipLaunchKernelGGL(kname<16>, grid, dim3(TPB), 0, (ctx)->stream, __VA_ARGS__); break; \
            default: hipLaunchKernelGGL(kname<8>, grid, dim3(TPB), 0, (ctx)->stream, __VA_ARGS__); break; \
        }                                                                                                \
    } while (0)

// Kernels whose PCG storage depends on the precision mode.
#define PEARL_LAUNCH_K(ctx, kname, MX, grid, ...)                                                             \
    do {                                                                                                      \
        switch ((ctx)->k) {                                                                                   \
            case 1: hipLaunchKernelGGL((kname<1, MX>), grid, dim3(TPB), 0, (ctx)->stream, __VA_ARGS__); break;   \
            case 2: hipLaunchKernelGGL((kname<2, MX>), grid, dim3(TPB), 0, (ctx)->stream, __VA_ARGS__); break;   \
            case 4: hipLaunchKernelGGL((kname<4, MX>), grid, dim3(TPB), 0, (ctx)->stream, __VA_ARGS__); break;   \
            case 16: hipLaunchKernelGGL((kname<16, MX>), grid, dim3(TPB), 0, (ctx)->stream, __VA_ARGS__); break; \
            default: hipLaunchKernelGGL((kname<8, MX>), grid, dim3(TPB), 0, (ctx)->stream, __VA_ARGS__); break;  \
        }                                                                                                     \
    } while (0)
#define PEARL_LAUNCH_P(ctx, kname, grid, ...)                          \
    do {                                                               \
        if ((ctx)->mx) PEARL_LAUNCH_K(ctx, kname, true, grid, __VA_ARGS__);  \
        else PEARL_LAUNCH_K(ctx, kname, false, grid, __VA_ARGS__);          \
    } while (0)

// cg_update runs TPB * cgb_split<k> threads per block (two per column).
void launch_cg_b(pearl_ctx* c, dim3 grid, const Bufs& B, int slot) {
    auto go = [&](auto ty, auto mx) {
        constexpr int TY = decltype(ty)::value;
        constexpr bool MX = decltype(mx)::value;
        hipLaunchKernelGGL((k_cg_b<TY, MX>), grid, dim3(TPB * cgb_split<TY>()), 0, c->stream, B, slot);
    };
    auto by_k = [&](auto mx) {
        switch (c->k) {
            case 1: go(std::integral_constant<int, 1>{}, mx); break;
            case 2: go(std::integral_constant<int, 2>{}, mx); break;
            case 4: go(std::integral_constant<int, 4>{}, mx); break;
            case 16: go(std::integral_constant<int, 16>{}, mx); break;
            default: go(std::integral_constant<int, 8>{}, mx); break;
        }
    };
    if (c->mx) by_k(std::true_type{});
    else by_k(std::false_type{});
}

// propose_phi in the context's form: FP32 stencil arithmetic only in the
// PEARL_FP32_STENCIL form.
void launch_propose(pearl_ctx* c, dim3 grid, const Bufs& B, int slot) {
    if (c->mx == PEARL_FP32_STENCIL) PEARL_LAUNCH_K(c, k_propose, true, grid, B, slot);
    else PEARL_LAUNCH_K(c, k_propose, false, grid, B, slot);
}

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

// Enqueue one round: [begin], [state], propose, prepare, slots x
// (cg_a, cg_b), finalize, [phase-only], [turn], [fixed-run copy].  `turn`
// decides the attempt and starts the next step or attempt (taken by the
// round's last finisher, turn_domain), so the round after it needs no begin: only the first round of a batch carries a
// leading begin, for the state the host left.  With `timed`, each cell
// kernel and its finisher are bracketed by events (the cell kernel alone,
// then the finisher) and the cell kernel writes its work record into
// kinds[slot].  The state reductions (e0, mass0 of a step) run only in a
// round whose flag asks for them: their cached values only turn valid
// during a run (host calls invalidate them), so a run needs them at most in
// its first round.  The phase-only kernels are present only when a domain of
// the context is phase-only.  A predicated-off launch costs several us.
void enqueue_round(pearl_ctx* c, int slots, bool timed, bool with_begin, bool with_turn, bool with_state,
                   bool with_save = false) {
    const dim3 g(static_cast<unsigned>(c->total_blocks));
    const dim3 gd((c->ndom + 63) / 64);
    int slot = 0;
    auto ev = [&](int kid, auto&& cell, auto&& fin) {
        if (timed) {
            hipEvent_t a = get_event(c), b = get_event(c), e = get_event(c);
            (void)hipEventRecord(a, c->stream);
            cell(slot);
            (void)hipEventRecord(b, c->stream);
            fin();
            (void)hipEventRecord(e, c->stream);
            c->pending.push_back({kid, slot, a, b});
            c->pending.push_back({PEARL_KERNEL_FINISH, -1, b, e});
        } else {
            cell(-1);
            fin();
        }
        ++slot;
    };
    // a cell kernel whose reduction needs no finisher (propose: atomic maxima
    // that the next kernel reads)
    auto ev_cell = [&](int kid, auto&& cell) {
        if (timed) {
            hipEvent_t a = get_event(c), b = get_event(c);
            (void)hipEventRecord(a, c->stream);
            cell(slot);
            (void)hipEventRecord(b, c->stream);
            c->pending.push_back({kid, slot, a, b});
        } else {
            cell(-1);
        }
        ++slot;
    };
    // finisher kernels: one block per domain
    const dim3 gf(static_cast<unsigned>(c->ndom));
    const dim3 tb(TPB);
    hipStream_t st = c->stream;
    const Bufs& B = c->B;
    if (with_begin) hipLaunchKernelGGL(k_begin, gd, dim3(64), 0, st, B);
    if (with_state)
        ev(PEARL_KERNEL_STATE_REDUCE, [&](int s) { PEARL_LAUNCH(c, k_state, g, B, s); },
           [&] { hipLaunchKernelGGL(k_state_fin, gf, tb, 0, st, B); });
    ev_cell(PEARL_KERNEL_PROPOSE, [&](int s) { launch_propose(c, g, B, s); });
    ev(PEARL_KERNEL_PREPARE, [&](int s) { PEARL_LAUNCH_P(c, k_prepare, g, B, s); },
       [&] { hipLaunchKernelGGL(k_prepare_fin, gf, tb, 0, st, B); });
    for (int j = 0; j < slots; ++j) {
        ev(PEARL_KERNEL_CG_APPLY, [&](int s) { PEARL_LAUNCH_P(c, k_cg_a, g, B, s); },
           [&] { hipLaunchKernelGGL(k_cg_a_fin, gf, tb, 0, st, B); });
        ev(PEARL_KERNEL_CG_UPDATE, [&](int s) { launch_cg_b(c, g, B, s); },
           [&] { hipLaunchKernelGGL(k_cg_b_fin, gf, tb, 0, st, B); });
    }
    // the round's last finisher also takes the turn when the round ends in
    // one
    const bool turn_fin = with_turn && !c->any_phase_only, turn_po = with_turn && c->any_phase_only;
    ev(PEARL_KERNEL_FINALIZE, [&](int s) { PEARL_LAUNCH_P(c, k_finalize, g, B, s); },
       [&] {
           if (turn_fin) hipLaunchKernelGGL(k_finalize_fin<true>, gf, tb, 0, st, B);
           else hipLaunchKernelGGL(k_finalize_fin<false>, gf, tb, 0, st, B);
       });
    if (c->any_phase_only)
        ev(PEARL_KERNEL_FINALIZE, [&](int s) { PEARL_LAUNCH(c, k_phase_only, g, B, s); },
           [&] {
               if (turn_po) hipLaunchKernelGGL(k_phase_only_fin<true>, gf, tb, 0, st, B);
               else hipLaunchKernelGGL(k_phase_only_fin<false>, gf, tb, 0, st, B);
           });
    if (with_save) {
        hipLaunchKernelGGL(k_fixed_save, dim3(static_cast<unsigned>(std::min<long long>(c->total_blocks, 2048))), tb, 0,
                           st, B);
        hipLaunchKernelGGL(k_fixed_saved, dim3(1), dim3(64), 0, st, B);
    }
}

// Attribute the timed launches since the last harvest (stream idle) and
// the per-domain work counters to the kernels.
void harvest(pearl_ctx* c) {
    if (!c->pending.empty()) {
        std::vector<int> kinds(MAX_SLOTS, 0);
        (void)hipMemcpy(kinds.data(), c->d_kinds, MAX_SLOTS * sizeof(int), hipMemcpyDeviceToHost);
        for (auto& e : c->pending) {
            int kid = e.kid;
            if (kid != PEARL_KERNEL_FINISH) {
                const int kind = e.slot < MAX_SLOTS ? kinds[e.slot] : 0;
                if (kind == 0) kid = PEARL_KERNEL_CG_IDLE;
                else if (kid == PEARL_KERNEL_CG_APPLY && kind == 2) kid = PEARL_KERNEL_TRUE_RESIDUAL;
                else if (kid == PEARL_KERNEL_CG_APPLY && kind == 3) kid = PEARL_KERNEL_CG_FIRST;
            }
            float t = 0;
            (void)hipEventElapsedTime(&t, e.a, e.b);
            c->ms[kid] += t;
            c->launches[kid]++;
        }
        // every event of the round once (the middle one is shared by a cell
        // kernel and its finisher)
        for (auto& e : c->pending) {
            if (e.kid == PEARL_KERNEL_FINISH) c->free_events.push_back(e.b);
            else {
                c->free_events.push_back(e.a);
                c->free_events.push_back(e.b);
            }
        }
        c->pending.clear();
    }
    // bytes from the device work counters (exact per domain)
    for (int d = 0; d < c->ndom; ++d) {
        for (int r = 0; r < RUN_COUNT; ++r) {
            const RunBytes& rb = kRunBytes[r];
            const unsigned now = c->hS[d].runs[r];
            unsigned& seen = c->runs_seen[static_cast<size_t>(d) * NKIND + r];
            c->bytes[rb.kid] += rb.bytes[c->mx ? 1 : 0] * c->kp[d].n * static_cast<double>(now - seen);
            if (!c->prof && rb.launch) c->launches[rb.kid] += now - seen;
            seen = now;
        }
    }
}

int read_scalars(pearl_ctx* c) {
    HIP_CHECK(c, hipMemcpyAsync(c->hS, c->d_S, c->ndom * sizeof(DevScal), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    HIP_CHECK(c, hipGetLastError());
    harvest(c);
    return PEARL_OK;
}

int write_scalars(pearl_ctx* c) {
    HIP_CHECK(c, hipMemcpyAsync(c->d_S, c->hS, c->ndom * sizeof(DevScal), hipMemcpyHostToDevice, c->stream));
    return PEARL_OK;
}

// Replay `rounds` rounds: from the captured graph, or eagerly with events
// (profiling; the work records are read after every round).
// `rounds` rounds; the first carries the state reductions when `state_first`;
// with `save` every round ends with the fixed-run state copy.
int run_rounds(pearl_ctx* c, int rounds, bool state_first, bool save = false) {
    if (c->prof) {
        for (int r = 0; r < rounds; ++r) {
            hipLaunchKernelGGL(k_clear_kinds, dim3(1), dim3(MAX_SLOTS), 0, c->stream, c->d_kinds, MAX_SLOTS);
            enqueue_round(c, c->cg_slots, true, r == 0, true, state_first && r == 0, save);
            if (r + 1 < rounds) {
                const int st = read_scalars(c);
                if (st) return st;
            }
        }
        return PEARL_OK;
    }
    const int first = save ? 4 : (state_first ? 2 : 1), rest = save ? 3 : 0;
    for (int v : {first, rest}) {
        if (c->graph_exec[v] && c->graph_slots[v] == c->cg_slots) continue;
        if (c->graph_exec[v]) (void)hipGraphExecDestroy(c->graph_exec[v]);
        if (c->graph[v]) (void)hipGraphDestroy(c->graph[v]);
        c->graph_exec[v] = nullptr;
        c->graph[v] = nullptr;
        HIP_CHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        enqueue_round(c, c->cg_slots, false, v == 1 || v == 2 || v == 4, true, v == 2, v == 3 || v == 4);
        HIP_CHECK(c, hipStreamEndCapture(c->stream, &c->graph[v]));
        HIP_CHECK(c, hipGraphInstantiate(&c->graph_exec[v], c->graph[v], nullptr, nullptr, 0));
        c->graph_slots[v] = c->cg_slots;
    }
    for (int r = 0; r < rounds; ++r) HIP_CHECK(c, hipGraphLaunch(c->graph_exec[r == 0 ? first : rest], c->stream));
    return PEARL_OK;
}

int status_of(pearl_ctx* c, int d) {
    const DevScal& h = c->hS[d];
    const int why = (h.why >= 0 && h.why < 5) ? h.why : 0;
    switch (h.status) {
        case ST_OK: return PEARL_OK;
        case ST_INVALID_TIME: return fail(c, PEARL_ERR_INVALID_TIME, "invalid advance time");
        case ST_DT_MIN: return fail(c, PEARL_ERR_DT_MIN, std::string("step rejected below dt_min: ") + kWhy[why]);
        case ST_RETRY_LIMIT: return fail(c, PEARL_ERR_RETRY_LIMIT, std::string("step retry limit: ") + kWhy[why]);
        case ST_MIXTURE: return fail(c, PEARL_ERR_MIXTURE, "Invalid mixture coefficients");
        case ST_PROJECTION: return fail(c, PEARL_ERR_PROJECTION, "Simplex projection failed");
        case ST_FIXED_SOLVE: return fail(c, PEARL_ERR_ARGUMENT, "fixed step: the linear solve did not converge");
        default: return fail(c, PEARL_ERR_DEVICE, "unknown device status");
    }
}

void info_of(const DevScal& h, pearl_step_info* info) {
    info->accepted_dt = h.i_dt;
    info->max_phase_change = h.i_change;
    info->linear_residual = h.i_residual;
    info->closure = h.i_closure;
    info->energy_change = h.i_energy_change;
    info->cg_iterations = h.i_cg;
    info->retries = h.i_retries;
    info->why = h.status == ST_OK ? -1 : h.why;
}

// Run every domain until it has taken max_steps[d] accepted steps or reached
// stop_time[d] (the step loop between two output times).  Rounds are
// replayed in batches between synchronisations; domains that finish idle.
// With `always_start` a domain starts one step even at or past stop_time,
// as a direct call of advance does (it then fails with "invalid advance
// time", :1121).
int run_domains(pearl_ctx* c, const double* stop_time, const long long* max_steps, bool always_start) {
    int st = read_scalars(c);
    if (st) return st;
    long long most = 0;
    for (int d = 0; d < c->ndom; ++d) {
        DevScal& h = c->hS[d];
        h.status = ST_OK;
        h.why = -1;
        h.taken = 0;
        h.target_steps = max_steps[d];
        h.stop_time = stop_time[d];
        h.stop_eps = c->stop_eps;
        h.time = c->sc[d].time;
        h.dt_state = c->sc[d].dt;
        h.step = c->sc[d].step;
        h.rejected = c->sc[d].rejected;
        h.cg_total = 0;
        h.cg_peak = 0;
        h.i_dt = h.i_change = h.i_residual = h.i_closure = h.i_energy_change = 0;
        h.i_cg = h.i_retries = 0;
        const bool work = max_steps[d] > 0 && (always_start || h.time < stop_time[d]);
        h.phase = work ? PH_START : PH_IDLE;
        if (work) most = std::max(most, max_steps[d]);
    }
    st = write_scalars(c);
    if (st) return st;
    // Rounds per synchronisation.  An accepted step takes at least one round
    // (more after a rejection or a solve longer than the round's PCG slots),
    // and a round replayed after every domain has finished costs ~20
    // predicated launches, more than a synchronisation.  So a batch is the
    // fewest steps any busy domain can still need: its remaining step
    // budget, and the steps to its stop time at its current dt (dt only
    // grows by accepted steps, up to dt_max).  Batches that end short of
    // the run's end cost one more synchronisation.
    auto rounds_left = [&]() {
        long long best = 0;
        for (int d = 0; d < c->ndom; ++d) {
            const DevScal& h = c->hS[d];
            if (h.phase == PH_IDLE) continue;
            long long need = std::max<long long>(h.target_steps - h.taken, 1);
            const double span = h.stop_time - h.stop_eps - h.time, dt = std::max(h.dt_state, c->prm[d].dt_max);
            if (span > 0 && dt > 0 && span / dt < static_cast<double>(need))
                need = std::max<long long>(1, static_cast<long long>(std::ceil(span / dt)));
            best = std::max(best, need);
        }
        return std::min<long long>(std::max<long long>(best, 1), 64);
    };
    // the state reductions are needed only if a starting domain has no
    // cached e0 / mass0 (k_begin's need_state); they then run in round one
    bool state_first = false;
    for (int d = 0; d < c->ndom; ++d) {
        const DevScal& h = c->hS[d];
        if (h.phase == PH_IDLE) continue;
        const KP& P = c->kp[d];
        if ((P.check_energy && !h.e0_valid) || (!P.phase_only && !h.m0_valid)) state_first = true;
    }
    long long batch = rounds_left();
    while (true) {
        st = run_rounds(c, static_cast<int>(batch), state_first);
        state_first = false;
        if (st) return st;
        st = read_scalars(c);
        if (st) return st;
        bool busy = false;
        for (int d = 0; d < c->ndom; ++d) busy = busy || c->hS[d].phase != PH_IDLE;
        if (!busy) break;
        batch = rounds_left();
    }
    int peak = 0;
    for (int d = 0; d < c->ndom; ++d) {
        const DevScal& h = c->hS[d];
        c->sc[d].time = h.time;
        c->sc[d].dt = h.dt_state;
        c->sc[d].step = h.step;
        c->sc[d].rejected = h.rejected;
        peak = std::max(peak, h.cg_peak);
    }
    // PCG slots per round: the largest solve (its true-residual check runs
    // inside finalize; a longer solve continues in the next round).
    if (peak > 0) c->cg_slots = std::min(24, std::max(3, peak));
    return PEARL_OK;
}

bool params_ok(const pearl_params* p) {
    if (!p) return false;
    if (p->nx < 3 || p->ny < 2 || !(p->dx > 0) || !(p->dy > 0)) return false;
    // the kernels index the three phase planes of a domain as a*n + i in
    // 32 bits: 3n stays below 2^31 (with room for the neighbour offsets)
    if (3 * static_cast<long long>(p->nx) * p->ny >= (1LL << 31) - (1LL << 20)) return false;
    for (int i = 0; i < 3; ++i)
        if (!(p->A[i] > 0) || !std::isfinite(p->A[i]) || !(p->tau[i] > 0)) return false;
    if (!(p->eps > 0) || !(p->sigma > 0) || !(p->Vm > 0) || p->triple < 0) return false;
    // the kernels' constant divisors (dx, dy, their squares, Vm, 2A) lie in
    // [2^-150, 2^150], the range the fast constant-divisor quotient assumes
    auto in = [](double v, double lo, double hi) { return v >= lo && v <= hi; };
    if (!in(p->dx, 0x1p-75, 0x1p75) || !in(p->dy, 0x1p-75, 0x1p75) || !in(p->Vm, 0x1p-150, 0x1p150)) return false;
    for (int i = 0; i < 3; ++i)
        if (!in(2.0 * p->A[i], 0x1p-150, 0x1p150)) return false;
    for (int i = 0; i < 3; ++i)
        if (!(p->pair_sigma[i] >= 0) || !std::isfinite(p->pair_sigma[i])) return false;
    return true;
}

// Parameter block of one domain: every derived constant uses the
// reference's own expression.
KP make_kp(const pearl_params* params, double mu_reference, int k) {
    KP P{};
    P.nx = params->nx;
    P.ny = params->ny;
    P.n = params->nx * params->ny;
    P.tiles_x = (P.nx + TPB - 1) / TPB;
    P.nb = P.tiles_x * ((P.ny + k - 1) / k);
    P.dx = params->dx;
    P.dy = params->dy;
    P.dx2 = params->dx * params->dx;
    P.dy2 = params->dy * params->dy;
    P.rdx = 1.0 / P.dx;
    P.rdy = 1.0 / P.dy;
    P.rdx2 = 1.0 / P.dx2;
    P.rdy2 = 1.0 / P.dy2;
    P.rVm = 1.0 / params->Vm;
    P.r3 = 1.0 / 3.0;
    for (int i = 0; i < 3; ++i) {
        P.A[i] = params->A[i];
        P.B[i] = params->B[i];
        P.C[i] = params->C[i];
        P.D[i] = params->D[i];
        P.chi[i] = 1.0 / (2.0 * params->A[i]);
        P.qpsi[i] = -0.25 / params->A[i];
        P.twoA[i] = 2.0 * params->A[i];
        P.rtwoA[i] = 1.0 / P.twoA[i];
        P.tau[i] = params->tau[i];
    }
    P.tau_min = params->tau[0];
    if (params->tau[1] < P.tau_min) P.tau_min = params->tau[1];
    if (params->tau[2] < P.tau_min) P.tau_min = params->tau[2];
    P.f_tau_min = static_cast<float>(P.tau_min);
    const double eps = params->eps, sigma = params->sigma;
    P.obs = 16 * sigma / (std::acos(-1.) * std::acos(-1.) * eps);
    P.tr = params->triple / eps;
    P.face_coef = 2 * eps * sigma;
    P.eps_sigma = eps * sigma;
    P.pair = 0;
    for (int k = 0; k < 3; ++k) {
        const double s = params->pair_sigma[k] > 0 ? params->pair_sigma[k] : sigma;
        if (s != sigma) P.pair = 1;
        P.sig_p[k] = s;
        P.obs_p[k] = 16 * s / (std::acos(-1.) * std::acos(-1.) * eps);
        P.fc_p[k] = 2 * eps * s;
        P.es_p[k] = eps * s;
    }
    P.eps = eps;
    P.f_obs = static_cast<float>(P.obs);
    P.f_tr = static_cast<float>(P.tr);
    for (int k = 0; k < 3; ++k) {
        P.f_obs_p[k] = static_cast<float>(P.obs_p[k]);
        P.f_fc_p[k] = static_cast<float>(P.fc_p[k]);
        P.f_tau[k] = static_cast<float>(params->tau[k]);
    }
    P.f_rVm = static_cast<float>(1.0 / params->Vm);
    P.f_rdx2 = static_cast<float>(P.rdx2);
    P.f_rdy2 = static_cast<float>(P.rdy2);
    P.f_eps = static_cast<float>(eps);
    P.sigma = sigma;
    P.Vm = params->Vm;
    auto psi = [&](int p, double mu) {
        const double d = mu - params->B[p];
        return std::fma(-0.25 / params->A[p], d * d, params->C[p]);
    };
    P.psiref[0] = psi(0, mu_reference) - psi(2, mu_reference);
    P.psiref[1] = psi(1, mu_reference) - psi(2, mu_reference);
    P.psiref[2] = 0;
    P.mu_ref = mu_reference;
    P.dt_max = params->dt_max;
    P.dt_min = params->dt_min;
    P.max_phase_change = params->max_phase_change;
    P.cg_rtol = params->cg_rtol;
    P.cg_atol = params->cg_atol;
    P.closure_tol = params->closure_tol;
    P.energy_rtol = params->energy_rtol;
    P.energy_atol = params->energy_atol;
    P.mass_floor = params->nx * params->ny * params->dx * params->dy * 1e-6;
    P.energy_scale = params->sigma * std::max(params->nx * params->dx, params->ny * params->dy);
    P.cg_max = params->cg_max;
    P.max_retries = params->max_retries;
    P.check_energy = params->check_energy;
    P.phase_only = params->phase_only;
    return P;
}

int ctx_check(pearl_ctx* c, int d) {
    if (!c || d < 0 || d >= c->ndom) return PEARL_ERR_ARGUMENT;
    (void)hipSetDevice(c->device);
    return PEARL_OK;
}

double* dom_phi(pearl_ctx* c, int d, int buf) { return c->B.phi[buf] + 3 * c->kp[d].off; }
double* dom_c(pearl_ctx* c, int d, int buf) { return c->B.c[buf] + c->kp[d].off; }
double* dom_mu(pearl_ctx* c, int d, int buf) { return c->B.mu[buf] + c->kp[d].off; }

// valid/mass/energy of domain d's current state.
int state_direct(pearl_ctx* c, int d, bool mixture_is_error, double* valid, double* mass, double* energy) {
    const KP& P = c->kp[d];
    PEARL_LAUNCH(c, k_state_direct, dim3(P.nb), c->B, d);
    hipLaunchKernelGGL(k_state_direct_fin, dim3(1), dim3(TPB), 0, c->stream, c->B, d);
    const int st = read_scalars(c);
    if (st) return st;
    const DevScal& h = c->hS[d];
    if (mixture_is_error && (h.err & ERRF_MIXTURE_STATE)) return fail(c, PEARL_ERR_MIXTURE, "Invalid mixture coefficients");
    if (valid) *valid = h.e_valid;
    if (mass) *mass = h.e_mass * P.dx * P.dy;
    if (energy) *energy = h.e_energy * P.dx * P.dy;
    return PEARL_OK;
}

}  // namespace

extern "C" {

int pearl_batch_create(int count, const pearl_params* params, const double* mu_references, int precision, int device,
                       int cells_per_thread, pearl_ctx** out) {
    pearl_ctx* c = nullptr;
    try {
        if (!out || count < 1 || !params || !mu_references) return PEARL_ERR_ARGUMENT;
        *out = nullptr;
        if (precision != PEARL_FP64 && precision != PEARL_MIXED && precision != PEARL_FP32_STENCIL)
            return PEARL_ERR_ARGUMENT;
        for (int d = 0; d < count; ++d)
            if (!params_ok(&params[d])) return PEARL_ERR_ARGUMENT;
        // rows per tile; 4 measured best on the 4096x2048 workload (2, 8, 16
        // within 2-7 % of it)
        int k = cells_per_thread;
        if (k <= 0) k = 4;
        if (k != 1 && k != 2 && k != 4 && k != 8 && k != 16) return PEARL_ERR_ARGUMENT;
        if (hipSetDevice(device) != hipSuccess) {
            (void)hipGetLastError();  // leave no error for the next call's launch check
            return PEARL_ERR_DEVICE;
        }
        c = new (std::nothrow) pearl_ctx();
        if (!c) return PEARL_ERR_ALLOC;
        c->device = device;
        c->ndom = count;
        c->k = k;
        c->mx = precision;
        c->prm.assign(params, params + count);
        for (int d = 0; d < count; ++d) c->any_phase_only = c->any_phase_only || params[d].phase_only != 0;
        c->sc.assign(count, pearl_scalars{});
        c->runs_seen.assign(static_cast<size_t>(count) * NKIND, 0u);
        auto bail0 = [&](int code) {
            delete c;
            return code;
        };
        // Pack the domains; each slice starts on a 64-cell (512-byte) boundary.
        long long off = 0, boff = 0, poff = 0;
        std::vector<int> blk2dom;
        for (int d = 0; d < count; ++d) {
            KP P = make_kp(&params[d], mu_references[d], k);
            P.off = off;
            P.boff = static_cast<int>(boff);
            P.poff = poff;
            off += (static_cast<long long>(P.n) + 63) / 64 * 64;
            boff += P.nb;
            // one launch grid holds every block, of up to 2 * TPB work-items
            // each, and a grid's work-items must stay below 2^32
            if (boff > 4294967295LL / (2 * TPB)) return bail0(PEARL_ERR_ARGUMENT);
            poff += 8LL * P.nb + 8;
            blk2dom.insert(blk2dom.end(), P.nb, d);
            c->kp.push_back(P);
        }
        c->total_cells = off;
        c->total_blocks = boff;
        c->total_parts = poff;
        // Planes are spaced by the packed cells plus 576 cells (4.5 KiB): with a
        // power-of-two domain the bare spacing would put the same cell of every
        // plane on the same cache set and memory channel.
        const size_t plane = static_cast<size_t>(off) + 576;
        const size_t total = plane * (3 * 2 + 2 + 2 + 10) + static_cast<size_t>(poff) + 64;
        auto bail = [&](int code) {
            pearl_destroy(c);
            (void)hipGetLastError();  // the failed HIP call's error is reported here only
            return code;
        };
        if (hipMalloc(&c->mem, total * sizeof(double)) != hipSuccess) return bail(PEARL_ERR_ALLOC);
        double* q = c->mem;
        for (int b = 0; b < 2; ++b) {
            c->B.phi[b] = q;
            q += 3 * plane;
        }
        for (int b = 0; b < 2; ++b) {
            c->B.c[b] = q;
            q += plane;
        }
        for (int b = 0; b < 2; ++b) {
            c->B.mu[b] = q;
            q += plane;
        }
        double** work[] = {&c->B.chi, &c->B.K, &c->B.rhs, &c->B.x, &c->B.diag,
                           &c->B.r,   &c->B.z,    &c->B.p0, &c->B.p1, &c->B.Ap};
        for (double** w : work) {
            *w = q;
            q += plane;
        }
        c->B.part = q;
        if (hipMalloc(&c->d_kp, count * sizeof(KP)) != hipSuccess) return bail(PEARL_ERR_ALLOC);
        if (hipMalloc(&c->d_S, count * sizeof(DevScal)) != hipSuccess) return bail(PEARL_ERR_ALLOC);
        if (hipMalloc(&c->d_blk2dom, blk2dom.size() * sizeof(int)) != hipSuccess) return bail(PEARL_ERR_ALLOC);
        if (hipMalloc(&c->d_kinds, MAX_SLOTS * sizeof(int)) != hipSuccess) return bail(PEARL_ERR_ALLOC);
        if (hipHostMalloc(&c->hS, count * sizeof(DevScal)) != hipSuccess) return bail(PEARL_ERR_ALLOC);
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return bail(PEARL_ERR_DEVICE);
        c->B.kp = c->d_kp;
        c->B.S = c->d_S;
        c->B.blk2dom = c->d_blk2dom;
        c->B.kinds = c->d_kinds;
        c->B.ndom = count;
        std::memset(c->hS, 0, count * sizeof(DevScal));
        for (int d = 0; d < count; ++d) c->hS[d].phase = PH_IDLE;
        if (hipMemcpyAsync(c->d_kp, c->kp.data(), count * sizeof(KP), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            hipMemcpyAsync(c->d_blk2dom, blk2dom.data(), blk2dom.size() * sizeof(int), hipMemcpyHostToDevice, c->stream) !=
                hipSuccess ||
            hipMemcpyAsync(c->d_S, c->hS, count * sizeof(DevScal), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            hipMemsetAsync(c->mem, 0, total * sizeof(double), c->stream) != hipSuccess ||
            hipMemsetAsync(c->d_kinds, 0, MAX_SLOTS * sizeof(int), c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            return bail(PEARL_ERR_DEVICE);
        *out = c;
        return PEARL_OK;
    } catch (const std::bad_alloc&) {
        if (c) pearl_destroy(c);  // the partly built context
        return alloc_fail(nullptr);
    }
}

int pearl_create(const pearl_params* params, double mu_reference, int precision, int device, int cells_per_thread,
                 pearl_ctx** out) {
    return pearl_batch_create(1, params, &mu_reference, precision, device, cells_per_thread, out);
}

int pearl_destroy(pearl_ctx* c) {
    if (!c) return PEARL_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& e : c->pending) {  // a finisher's first event is its cell kernel's second
        if (e.kid != PEARL_KERNEL_FINISH) (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    for (auto e : c->free_events) (void)hipEventDestroy(e);
    for (int v = 0; v < 5; ++v) {
        if (c->graph_exec[v]) (void)hipGraphExecDestroy(c->graph_exec[v]);
        if (c->graph[v]) (void)hipGraphDestroy(c->graph[v]);
    }
    if (c->mem) (void)hipFree(c->mem);
    if (c->d_dt1) (void)hipFree(c->d_dt1);
    if (c->d_dtn) (void)hipFree(c->d_dtn);
    if (c->d_kp) (void)hipFree(c->d_kp);
    if (c->d_S) (void)hipFree(c->d_S);
    if (c->d_blk2dom) (void)hipFree(c->d_blk2dom);
    if (c->d_kinds) (void)hipFree(c->d_kinds);
    if (c->hS) (void)hipHostFree(c->hS);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return PEARL_OK;
}

const char* pearl_last_error(const pearl_ctx* c) { return c ? c->err.c_str() : "null context"; }

int pearl_batch_size(const pearl_ctx* c, int* count) {
    if (!c || !count) return PEARL_ERR_ARGUMENT;
    *count = c->ndom;
    return PEARL_OK;
}

int pearl_batch_set_state(pearl_ctx* c, int d, const double* phi, const double* carbon, const double* mu,
                          const pearl_scalars* scalars) {
    int st = ctx_check(c, d);
    if (st) return st;
    st = read_scalars(c);
    if (st) return st;
    DevScal& h = c->hS[d];
    const size_t n = static_cast<size_t>(c->kp[d].n);
    if (phi) HIP_CHECK(c, hipMemcpyAsync(dom_phi(c, d, h.cur), phi, 3 * n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (carbon) HIP_CHECK(c, hipMemcpyAsync(dom_c(c, d, h.cur), carbon, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (mu) HIP_CHECK(c, hipMemcpyAsync(dom_mu(c, d, h.cur), mu, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (scalars) c->sc[d] = *scalars;
    h.time = c->sc[d].time;
    h.dt_state = c->sc[d].dt;
    h.step = c->sc[d].step;
    h.rejected = c->sc[d].rejected;
    // the cached e0 / mass0 belong to the previous fields
    if (phi || carbon || mu) {
        h.e0_valid = 0;
        h.m0_valid = 0;
    }
    h.err = 0;
    h.status = ST_OK;
    h.phase = PH_IDLE;
    st = write_scalars(c);
    if (st) return st;
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return PEARL_OK;
}

int pearl_batch_get_state(pearl_ctx* c, int d, double* phi, double* carbon, double* mu, pearl_scalars* scalars) {
    int st = ctx_check(c, d);
    if (st) return st;
    st = read_scalars(c);
    if (st) return st;
    const int cur = c->hS[d].cur;
    const size_t n = static_cast<size_t>(c->kp[d].n);
    if (phi) HIP_CHECK(c, hipMemcpyAsync(phi, dom_phi(c, d, cur), 3 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (carbon) HIP_CHECK(c, hipMemcpyAsync(carbon, dom_c(c, d, cur), n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (mu) HIP_CHECK(c, hipMemcpyAsync(mu, dom_mu(c, d, cur), n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    if (scalars) *scalars = c->sc[d];
    return PEARL_OK;
}

int pearl_set_state(pearl_ctx* c, const double* phi, const double* carbon, const double* mu, const pearl_scalars* s) {
    return pearl_batch_set_state(c, 0, phi, carbon, mu, s);
}

int pearl_get_state(pearl_ctx* c, double* phi, double* carbon, double* mu, pearl_scalars* s) {
    return pearl_batch_get_state(c, 0, phi, carbon, mu, s);
}

namespace {
int batch_run(pearl_ctx* c, const double* stop_times, const long long* max_steps, long long* taken,
              long long* cg_total, int* cg_peak, pearl_step_info* last, int* status, bool always_start) {
    int st = ctx_check(c, 0);
    if (st) return st;
    if (!stop_times || !max_steps) return PEARL_ERR_ARGUMENT;
    st = run_domains(c, stop_times, max_steps, always_start);
    if (st) return st;
    int first = PEARL_OK, first_d = -1;
    for (int d = 0; d < c->ndom; ++d) {
        const DevScal& h = c->hS[d];
        if (taken) taken[d] = h.taken;
        if (cg_total) cg_total[d] += h.cg_total;
        if (cg_peak) cg_peak[d] = std::max(cg_peak[d], h.cg_peak);
        if (last) info_of(h, &last[d]);
        const int sd = status_of(c, d);
        if (status) status[d] = sd;
        if (sd && first_d < 0) {
            first = sd;
            first_d = d;
        }
    }
    if (first_d >= 0) (void)status_of(c, first_d);  // the message of the first failing domain
    return first;
}

int single_run(pearl_ctx* c, double stop_time, long long max_steps, long long* taken, long long* cg_total,
               int* cg_peak, pearl_step_info* last, bool always_start) {
    int status = 0;
    pearl_step_info info{};
    long long t = 0, cgt = cg_total ? *cg_total : 0;
    int cgp = cg_peak ? *cg_peak : 0;
    const int st = batch_run(c, &stop_time, &max_steps, &t, &cgt, &cgp, &info, &status, always_start);
    if (taken) *taken = t;
    if (cg_total) *cg_total = cgt;
    if (cg_peak) *cg_peak = cgp;
    if (last) *last = info;
    return st;
}
}  // namespace

int pearl_batch_run(pearl_ctx* c, const double* stop_times, const long long* max_steps, long long* taken,
                    long long* cg_total, int* cg_peak, pearl_step_info* last, int* status) {
    try {
        return batch_run(c, stop_times, max_steps, taken, cg_total, cg_peak, last, status, false);
    } catch (const std::bad_alloc&) {
        return alloc_fail(c);
    }
}

int pearl_run(pearl_ctx* c, double stop_time, long long max_steps, long long* taken, long long* cg_total, int* cg_peak,
              pearl_step_info* last) {
    try {
        int status = 0;
        pearl_step_info info{};
        long long t = 0, cgt = cg_total ? *cg_total : 0;
        int cgp = cg_peak ? *cg_peak : 0;
        const int st = batch_run(c, &stop_time, &max_steps, &t, &cgt, &cgp, &info, &status, false);
        if (taken) *taken = t;
        if (cg_total) *cg_total = cgt;
        if (cg_peak) *cg_peak = cgp;
        if (last) *last = info;
        return st;
    } catch (const std::bad_alloc&) {
        return alloc_fail(c);
    }
}

int pearl_set_stop_tolerance(pearl_ctx* c, double eps) {
    if (!c) return PEARL_ERR_ARGUMENT;
    if (!(eps >= 0) || !std::isfinite(eps)) return fail(c, PEARL_ERR_ARGUMENT, "invalid stop tolerance");
    c->stop_eps = eps;
    return PEARL_OK;
}

int pearl_advance(pearl_ctx* c, double stop_time, pearl_step_info* info) {
    try {
        if (!c || !info) return PEARL_ERR_ARGUMENT;
        long long taken = 0;
        std::memset(info, 0, sizeof(*info));
        return single_run(c, stop_time, 1, &taken, nullptr, nullptr, info, true);
    } catch (const std::bad_alloc&) {
        return alloc_fail(c);
    }
}

int pearl_maybe_shift(pearl_ctx* c, int moving_window, double trigger, double target, double c_gamma, int* shifted,
                      double* front) {
    try {
        int st = ctx_check(c, 0);
        if (st) return st;
        if (!shifted) return PEARL_ERR_ARGUMENT;
        *shifted = 0;
        if (front) *front = -1;
        if (!moving_window) return PEARL_OK;
        const KP& P = c->kp[0];
        PEARL_LAUNCH(c, k_front, dim3(P.nb), c->B, 0);
        hipLaunchKernelGGL(k_front_fin, dim3(1), dim3(TPB), 0, c->stream, c->B, 0);
        st = read_scalars(c);
        if (st) return st;
        const double fr = c->hS[0].front;
        if (front) *front = fr;
        if (fr < trigger) return PEARL_OK;
        const int rows = static_cast<int>(std::floor((fr - target) / P.dy));
        const int scan_rows = std::max(0, std::min(rows, P.ny));
        double removed = 0;
        if (scan_rows > 0) {
            const int nbs = P.tiles_x * ((scan_rows + c->k - 1) / c->k);
            PEARL_LAUNCH(c, k_strip, dim3(nbs), c->B, 0, scan_rows, nbs);
            hipLaunchKernelGGL(k_strip_fin, dim3(1), dim3(TPB), 0, c->stream, c->B, 0, nbs);
            st = read_scalars(c);
            if (st) return st;
            if (c->hS[0].gmax > 1e-6) return fail(c, PEARL_ERR_UNSAFE_SHIFT, "unsafe shift: removed strip contains gamma");
            removed = c->hS[0].removed;
        }
        if (rows <= 0 || rows >= P.ny || c_gamma < 0 || c_gamma > 1) return fail(c, PEARL_ERR_INVALID_SHIFT, "invalid shift");
        const double added = rows * P.nx * c_gamma * P.dx * P.dy;
        // the flags k_shift raises start clear: an earlier failed shift
        // leaves them set
        hipLaunchKernelGGL(k_clear_err, dim3(1), dim3(64), 0, c->stream, c->B, 0,
                           static_cast<int>(ERRF_MIXTURE | ERRF_NONFINITE_CARBON));
        PEARL_LAUNCH(c, k_shift, dim3(P.nb), c->B, 0, rows, c_gamma);
        st = read_scalars(c);
        if (st) return st;
        DevScal& h = c->hS[0];
        if (h.err & ERRF_MIXTURE) return fail(c, PEARL_ERR_MIXTURE, "Invalid mixture coefficients");
        if (h.err & ERRF_NONFINITE_CARBON) return fail(c, PEARL_ERR_NONFINITE_CARBON, "Nonfinite total carbon");
        h.cur ^= 1;
        h.e0_valid = 0;
        h.m0_valid = 0;
        st = write_scalars(c);
        if (st) return st;
        c->sc[0].shift_distance += rows * P.dy;
        c->sc[0].shifts++;
        c->sc[0].carbon_added += added;
        c->sc[0].carbon_removed += removed;
        *shifted = 1;
        HIP_CHECK(c, hipStreamSynchronize(c->stream));
        return PEARL_OK;
    } catch (const std::bad_alloc&) {
        return alloc_fail(c);
    }
}

int pearl_diagnostics(pearl_ctx* c, pearl_diag* out) {
    try {
        int st = ctx_check(c, 0);
        if (st) return st;
        if (!out) return PEARL_ERR_ARGUMENT;
        const KP& P = c->kp[0];
        PEARL_LAUNCH(c, k_diagnostics, dim3(P.nb), c->B, 0);
        hipLaunchKernelGGL(k_diagnostics_fin, dim3(1), dim3(TPB), 0, c->stream, c->B, 0);
        st = read_scalars(c);
        if (st) return st;
        const DevScal& h = c->hS[0];
        if (!c->prm[0].phase_only && (h.err & ERRF_MIXTURE_STATE))
            return fail(c, PEARL_ERR_MIXTURE, "Invalid mixture coefficients");
        out->c_min = h.dmin_c;
        out->c_max = h.dmax_c;
        out->mu_min = h.dmin_mu;
        out->mu_max = h.dmax_mu;
        out->phi_sum_error = h.dsumerr;
        out->mass = h.dmass * P.dx * P.dy;
        out->energy = h.denergy * P.dx * P.dy;
        return PEARL_OK;
    } catch (const std::bad_alloc&) {
        return alloc_fail(c);
    }
}

int pearl_propose_phi(pearl_ctx* c, double dt, double* phi_out, double* max_change) {
    try {
        int st = ctx_check(c, 0);
        if (st) return st;
        if (!max_change) return PEARL_ERR_ARGUMENT;
        st = read_scalars(c);
        if (st) return st;
        DevScal& h = c->hS[0];
        h.phase = PH_PROPOSE;
        h.need_state = 0;
        h.dt = dt;
        h.status = ST_OK;
        for (auto& b : h.chg) b = 0;  // propose's reduction slots (the device's begin clears them in a round)
        h.proj = 0;
        st = write_scalars(c);
        if (st) return st;
        launch_propose(c, dim3(static_cast<unsigned>(c->total_blocks)), c->B, -1);
        hipLaunchKernelGGL(k_propose_fin, dim3(static_cast<unsigned>(c->ndom)), dim3(TPB), 0, c->stream, c->B);
        st = read_scalars(c);
        if (st) return st;
        const int status = h.status;
        *max_change = h.change;
        h.phase = PH_IDLE;
        h.status = ST_OK;
        st = write_scalars(c);
        if (st) return st;
        if (status == ST_PROJECTION) return fail(c, PEARL_ERR_PROJECTION, "Simplex projection failed");
        if (phi_out) {
            HIP_CHECK(c, hipMemcpyAsync(phi_out, dom_phi(c, 0, h.cur ^ 1), 3 * static_cast<size_t>(c->kp[0].n) * sizeof(double),
                                        hipMemcpyDeviceToHost, c->stream));
        }
        HIP_CHECK(c, hipStreamSynchronize(c->stream));
        return PEARL_OK;
    } catch (const std::bad_alloc&) {
        return alloc_fail(c);
    }
}

int pearl_implicit_diffusion(pearl_ctx* c, const double* phi_trial, double dt, double* carbon_out, double* mu_out,
                             pearl_diffusion_info* info) {
    try {
        int st = ctx_check(c, 0);
        if (st) return st;
        if (!phi_trial || !info || c->prm[0].phase_only) return PEARL_ERR_ARGUMENT;
        st = read_scalars(c);
        if (st) return st;
        DevScal& h = c->hS[0];
        const size_t n = static_cast<size_t>(c->kp[0].n);
        HIP_CHECK(c, hipMemcpyAsync(dom_phi(c, 0, h.cur ^ 1), phi_trial, 3 * n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        h.phase = PH_PREPARE;
        h.live = 1;
        h.cg_alone = 0;
        h.cg_state = CG_IDLE;
        h.cg_it = 0;
        h.cg_res = 0;
        h.closure = 0;
        h.pad = 0;
        h.dt = dt;
        h.status = ST_OK;
        h.err = 0;
        st = write_scalars(c);
        if (st) return st;
        // rounds without begin/decide until the solve leaves PH_SOLVE
        for (int guard = 0; guard < 100000; ++guard) {
            if (c->prof) {
                hipLaunchKernelGGL(k_clear_kinds, dim3(1), dim3(MAX_SLOTS), 0, c->stream, c->d_kinds, MAX_SLOTS);
                enqueue_round(c, c->cg_slots, true, false, false, false);
            } else {
                enqueue_round(c, c->cg_slots, false, false, false, false);
            }
            st = read_scalars(c);
            if (st) return st;
            if (h.phase != PH_SOLVE) break;
        }
        if (h.status != ST_OK) {
            const int code = status_of(c, 0);
            h.phase = PH_IDLE;
            h.status = ST_OK;
            (void)write_scalars(c);
            return code;
        }
        const bool finalized = h.cg_state == CG_OK;
        info->it = h.cg_it;
        info->residual = h.cg_res;
        info->closure = finalized ? h.closure : 0;
        info->ok = finalized && info->closure <= c->prm[0].closure_tol;
        h.phase = PH_IDLE;
        st = write_scalars(c);
        if (st) return st;
        const int src = finalized ? (h.cur ^ 1) : h.cur;
        if (carbon_out) HIP_CHECK(c, hipMemcpyAsync(carbon_out, dom_c(c, 0, src), n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        if (mu_out) HIP_CHECK(c, hipMemcpyAsync(mu_out, dom_mu(c, 0, src), n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(c, hipStreamSynchronize(c->stream));
        return PEARL_OK;
    } catch (const std::bad_alloc&) {
        return alloc_fail(c);
    }
}

namespace {
// `n` steps of the sizes dts[0..n-1] (a device array) that are always
// taken: propose_phi (:961-1002), then implicit_diffusion (:1010-1094) on the
// trial phases, then the trial state becomes current.  No acceptance test,
// no dt control: the differentiable map that the adjoint inverts.  Domain
// 0, coupled mode.  The steps run under the device controller in fixed-dt
// mode, from replayed rounds, with a synchronisation per batch of rounds.
// With `save` (device memory for n states of 5 planes), the state after
// step k is copied to save + (k-1)*5n as the run goes.
int run_fixed(pearl_ctx* c, const double* dts, int n, double* save = nullptr) {
    int st = read_scalars(c);
    if (st) return st;
    DevScal& h = c->hS[0];
    h.status = ST_OK;
    h.why = -1;
    h.taken = 0;
    h.target_steps = n;
    h.fixed = 1;
    h.fixed_dts = dts;
    h.fixed_save = save;
    h.fixed_saved = 0;
    h.cg_peak = 0;
    h.phase = PH_START;
    st = write_scalars(c);
    // Whatever happens below, the context leaves fixed-dt mode: the next
    // advance must run under the adaptive controller.
    struct Leave {
        pearl_ctx* c;
        ~Leave() {
            DevScal& h = c->hS[0];
            h.fixed = 0;
            h.fixed_dts = nullptr;
            h.fixed_save = nullptr;
            h.phase = PH_IDLE;
            (void)write_scalars(c);
        }
    } leave{c};
    if (st) return st;
    long long batch = std::min(n, 64);
    while (true) {
        st = run_rounds(c, static_cast<int>(batch), false, save != nullptr);
        if (st) return st;
        st = read_scalars(c);
        if (st) return st;
        if (h.phase == PH_IDLE) break;
        batch = std::max<long long>(1, std::min<long long>(64, h.target_steps - h.taken));
    }
    const int code = h.status == ST_OK ? PEARL_OK : status_of(c, 0);
    if (h.cg_peak > 0) c->cg_slots = std::min(24, std::max(3, h.cg_peak));
    h.status = ST_OK;
    c->sc[0].time = h.time;
    c->sc[0].step = h.step;
    return code;
}

int step_fixed(pearl_ctx* c, double dt) {
    if (!c->d_dt1 && hipMalloc(&c->d_dt1, sizeof(double)) != hipSuccess)
        return fail(c, PEARL_ERR_ALLOC, "out of device memory");
    HIP_CHECK(c, hipMemcpyAsync(c->d_dt1, &dt, sizeof(double), hipMemcpyHostToDevice, c->stream));
    return run_fixed(c, c->d_dt1, 1);
}
}  // namespace

int pearl_step_fixed(pearl_ctx* c, double dt) {
    try {
        int st = ctx_check(c, 0);
        if (st) return st;
        if (c->prm[0].phase_only || c->mx || !(dt > 0) || !std::isfinite(dt)) return PEARL_ERR_ARGUMENT;
        return step_fixed(c, dt);
    } catch (const std::bad_alloc&) {
        return alloc_fail(c);
    }
}

int pearl_run_fixed(pearl_ctx* c, int steps, const double* dts) {
    try {
        int st = ctx_check(c, 0);
        if (st) return st;
        if (c->ndom != 1 || c->prm[0].phase_only || c->mx || steps < 1 || !dts) return PEARL_ERR_ARGUMENT;
        for (int k = 0; k < steps; ++k)
            if (!(dts[k] > 0) || !std::isfinite(dts[k])) return PEARL_ERR_ARGUMENT;
        if (steps > c->d_dtn_cap) {
            if (c->d_dtn) (void)hipFree(c->d_dtn);
            c->d_dtn = nullptr;
            c->d_dtn_cap = 0;
            if (hipMalloc(&c->d_dtn, steps * sizeof(double)) != hipSuccess)
                return fail(c, PEARL_ERR_ALLOC, "out of device memory");
            c->d_dtn_cap = steps;
        }
        HIP_CHECK(c, hipMemcpyAsync(c->d_dtn, dts, steps * sizeof(double), hipMemcpyHostToDevice, c->stream));
        return run_fixed(c, c->d_dtn, steps);
    } catch (const std::bad_alloc&) {
        return alloc_fail(c);
    }
}

namespace {
// Device copy of the five state planes (phi x3, c, mu) of domain 0.
int copy_state(pearl_ctx* c, double* dst_phi, double* dst_c, double* dst_mu, const double* src_phi,
               const double* src_c, const double* src_mu) {
    const size_t n = static_cast<size_t>(c->kp[0].n) * sizeof(double);
    HIP_CHECK(c, hipMemcpyAsync(dst_phi, src_phi, 3 * n, hipMemcpyDeviceToDevice, c->stream));
    HIP_CHECK(c, hipMemcpyAsync(dst_c, src_c, n, hipMemcpyDeviceToDevice, c->stream));
    HIP_CHECK(c, hipMemcpyAsync(dst_mu, src_mu, n, hipMemcpyDeviceToDevice, c->stream));
    return PEARL_OK;
}

// Adjoint PCG solve A lam = gx (A of the trial phases, dt of the step),
// lam left in B.x.  Reuses the forward PCG kernels and their controller.
// `slots`: PCG slots enqueued per synchronisation; raised to the solve's
// iterations + 1 when a solve needs more (the next solve then fits).
int adjoint_solve(pearl_ctx* c, const Adj& A, int& slots) {
    DevScal& h = c->hS[0];
    h.phase = PH_SOLVE;
    h.cg_alone = 1;
    h.cg_state = CG_IDLE;
    h.cg_it = 0;
    h.cg_res = 0;
    h.pad = 0;
    h.pidx = 0;
    h.dt = A.dt;
    h.status = ST_OK;
    int st = write_scalars(c);
    if (st) return st;
    const KP& P = c->kp[0];
    const dim3 grid(static_cast<unsigned>(P.nb)), one(1), tb(TPB);
    PEARL_LAUNCH(c, k_adj_prepare, grid, c->B, A);
    hipLaunchKernelGGL(k_adj_prepare_fin, one, tb, 0, c->stream, c->B);
    // PCG slots, then one synchronisation to see whether the solve is done
    // (slots after convergence are predicated off)
    for (int guard = 0; guard < 10000; ++guard) {
        for (int k = 0; k < slots; ++k) {
            PEARL_LAUNCH_P(c, k_cg_a, grid, c->B, -1);
            hipLaunchKernelGGL(k_cg_a_fin, one, tb, 0, c->stream, c->B);
            launch_cg_b(c, grid, c->B, -1);
            hipLaunchKernelGGL(k_cg_b_fin, one, tb, 0, c->stream, c->B);
        }
        st = read_scalars(c);
        if (st) return st;
        if (h.cg_state == CG_OK) break;
        if (h.cg_state == CG_FAIL) return fail(c, PEARL_ERR_ARGUMENT, "adjoint: the linear solve did not converge");
    }
    slots = std::min(24, std::max(slots, h.cg_it + 1));
    h.phase = PH_IDLE;
    return write_scalars(c);
}
}  // namespace

int pearl_adjoint(pearl_ctx* c, int steps, const double* dts, int every, const double* phi_target, double w_front,
                  double front_target, double* loss, double* grad) {
    try {
        int st = ctx_check(c, 0);
        if (st) return st;
        if (c->ndom != 1 || c->mx || c->prm[0].phase_only || steps < 1 || every < 1 || !dts || !phi_target || !loss ||
            !grad || !std::isfinite(w_front) || w_front < 0)
            return PEARL_ERR_ARGUMENT;
        for (int k = 0; k < steps; ++k)
            if (!(dts[k] > 0) || !std::isfinite(dts[k])) return PEARL_ERR_ARGUMENT;
        const KP& P = c->kp[0];
        const size_t n = static_cast<size_t>(P.n);
        const int segs = (steps + every - 1) / every;
        const int seg_len = std::min(every, steps);
        // device memory: checkpoints, the states of one segment, the final state,
        // two gradient sets, the trial-phase gradient, v and the accumulators
        const size_t state = 5 * n;
        const size_t total =
            state * (static_cast<size_t>(segs) + seg_len + 1 + 1) + 2 * state + 6 * n + ADJ_THETA + static_cast<size_t>(steps);
        double* mem = nullptr;
        if (hipMalloc(&mem, total * sizeof(double)) != hipSuccess) return fail(c, PEARL_ERR_ALLOC, "adjoint: out of memory");
        struct Free {
            double* p;
            ~Free() { (void)hipFree(p); }
        } guard{mem};
        double* ck = mem;
        double* seg = ck + state * segs;
        double* fin = seg + state * (seg_len + 1);
        double* G[2] = {fin + state, fin + 2 * state};
        double* gphit = G[1] + state;
        double* vbuf = gphit + 3 * n;
        double* theta = vbuf + 3 * n;
        double* d_dts = theta + ADJ_THETA;
        auto ph = [&](double* s5) { return s5; };
        auto cc = [&](double* s5) { return s5 + 3 * n; };
        auto mm = [&](double* s5) { return s5 + 4 * n; };
        st = read_scalars(c);
        if (st) return st;
        DevScal& h = c->hS[0];
        auto save = [&](double* s5) {
            return copy_state(c, ph(s5), cc(s5), mm(s5), dom_phi(c, 0, h.cur), dom_c(c, 0, h.cur), dom_mu(c, 0, h.cur));
        };
        auto load = [&](double* s5) {
            return copy_state(c, dom_phi(c, 0, h.cur), dom_c(c, 0, h.cur), dom_mu(c, 0, h.cur), ph(s5), cc(s5), mm(s5));
        };
        HIP_CHECK(c, hipMemcpyAsync(d_dts, dts, steps * sizeof(double), hipMemcpyHostToDevice, c->stream));
        // forward with checkpoints: one device-controlled run per segment
        if ((st = save(ck))) return st;
        for (int sgi = 0; sgi < segs; ++sgi) {
            const int s0 = sgi * every, len = std::min(every, steps - s0);
            if ((st = run_fixed(c, d_dts + s0, len))) return st;
            if (s0 + len < steps && (st = save(ck + state * (sgi + 1)))) return st;
        }
        if ((st = save(fin))) return st;
        const double t_end = h.time;
        const long long step_end = h.step;
        // loss and its gradient w.r.t. the final state (host)
        std::vector<double> phiN(3 * n), g(5 * n, 0.0);
        HIP_CHECK(c, hipMemcpyAsync(phiN.data(), ph(fin), 3 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(c, hipStreamSynchronize(c->stream));
        // compensated (Neumaier) sums: the loss and the front height do not
        // depend on the summation order beyond the last bits
        struct Sum {
            double s = 0, c = 0;
            void add(double v) {
                const double t = s + v;
                c += std::fabs(s) >= std::fabs(v) ? (s - t) + v : (v - t) + s;
                s = t;
            }
            double value() const { return s + c; }
        } l2s, hs;
        for (size_t i = 0; i < 3 * n; ++i) {
            const double d = phiN[i] - phi_target[i];
            l2s.add(d * d);
            g[i] = d;
        }
        for (size_t i = 0; i < n; ++i) hs.add(1.0 - phiN[2 * n + i]);
        const double l2 = l2s.value();
        const double height = hs.value() * P.dy / P.nx;
        const double dh_ = height - front_target;
        *loss = 0.5 * l2 + 0.5 * w_front * dh_ * dh_;
        for (size_t i = 0; i < n; ++i) g[2 * n + i] -= w_front * dh_ * P.dy / P.nx;
        HIP_CHECK(c, hipMemcpyAsync(G[1], g.data(), 5 * n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIP_CHECK(c, hipMemsetAsync(theta, 0, ADJ_THETA * sizeof(double), c->stream));
        // reverse sweep, one checkpointed segment at a time
        const dim3 grid(static_cast<unsigned>(P.nb)), one(1), tb(TPB);
        int cur = 1;  // G[cur] holds the gradient w.r.t. the later state
        int slots = c->cg_slots;
        for (int sgi = segs - 1; sgi >= 0; --sgi) {
            const int s0 = sgi * every, len = std::min(every, steps - s0);
            if ((st = load(ck + state * sgi))) return st;
            if ((st = save(seg))) return st;
            if ((st = run_fixed(c, d_dts + s0, len, seg + state))) return st;
            for (int k = len - 1; k >= 0; --k) {
                double* s_n = seg + state * k;
                double* s_n1 = seg + state * (k + 1);
                double* gin = G[cur];
                double* gout = G[cur ^ 1];
                Adj A{ph(s_n), mm(s_n), ph(s_n1), mm(s_n1), ph(gin), cc(gin), mm(gin), ph(gout), cc(gout), mm(gout),
                      gphit, vbuf, theta, dts[s0 + k]};
                if ((st = adjoint_solve(c, A, slots))) return st;
                PEARL_LAUNCH(c, k_adj_diffgrad, grid, c->B, A);
                hipLaunchKernelGGL(k_adj_diffgrad_fin, one, tb, 0, c->stream, c->B, A);
                PEARL_LAUNCH(c, k_adj_propose_a, grid, c->B, A);
                hipLaunchKernelGGL(k_adj_propose_a_fin, one, tb, 0, c->stream, c->B, A);
                PEARL_LAUNCH(c, k_adj_propose_b, grid, c->B, A);
                cur ^= 1;
            }
        }
        double th[ADJ_THETA];
        HIP_CHECK(c, hipMemcpyAsync(th, theta, ADJ_THETA * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        // the context ends on the forward's final state
        if ((st = load(fin))) return st;
        HIP_CHECK(c, hipStreamSynchronize(c->stream));
        HIP_CHECK(c, hipGetLastError());
        // grad: sigma (the three pair energies moved together), tau_at, tau_ag,
        // tau_tg, D_alpha, D_theta, D_gamma, then sigma_at, sigma_ag, sigma_tg
        grad[0] = th[0] + th[1] + th[2];
        for (int q = 0; q < 6; ++q) grad[1 + q] = th[3 + q];
        for (int q = 0; q < 3; ++q) grad[7 + q] = th[q];
        st = read_scalars(c);
        if (st) return st;
        h.time = t_end;
        h.step = step_end;
        h.phase = PH_IDLE;
        h.e0_valid = 0;
        h.m0_valid = 0;
        st = write_scalars(c);
        if (st) return st;
        c->sc[0].time = t_end;
        c->sc[0].step = step_end;
        HIP_CHECK(c, hipStreamSynchronize(c->stream));
        return PEARL_OK;
    } catch (const std::bad_alloc&) {
        return alloc_fail(c);
    }
}

int pearl_smooth(pearl_ctx* c, int max_steps, int min_steps, double energy_rtol, double max_phase_change,
                 pearl_smooth_report* rep) {
    try {
        int st = ctx_check(c, 0);
        if (st) return st;
        if (!rep) return PEARL_ERR_ARGUMENT;
        std::memset(rep, 0, sizeof(*rep));
        if (max_steps < 0 || min_steps < 0 || min_steps > max_steps || !(energy_rtol > 0) || !(max_phase_change > 0) ||
            max_phase_change > .2)
            return fail(c, PEARL_ERR_ARGUMENT, "invalid sharp-smoothing controls");
        st = read_scalars(c);
        if (st) return st;
        const KP& P = c->kp[0];
        const pearl_params& prm = c->prm[0];
        const dim3 grid(static_cast<unsigned>(P.nb)), one(1), tb(TPB);
        DevScal& h = c->hS[0];
        const int cur0 = h.cur;
        // interface energy of the current phases (:880)
        PEARL_LAUNCH(c, k_iface_energy, grid, c->B, 0);
        hipLaunchKernelGGL(k_iface_energy_fin, one, tb, 0, c->stream, c->B, 0);
        st = read_scalars(c);
        if (st) return st;
        rep->energy_initial = h.sm_energy * P.dx * P.dy;
        if (max_steps == 0) rep->converged = 1;  // the energy alone
        double previous = rep->energy_initial;
        const double eps = prm.eps, sigma = prm.sigma;
        const double base = .12 * std::min(prm.dx * prm.dx, prm.dy * prm.dy) / (eps * sigma);  // :884
        int result = PEARL_OK;
        for (int step = 1; step <= max_steps; ++step) {
            PEARL_LAUNCH(c, k_smooth_force, grid, c->B, 0);
            hipLaunchKernelGGL(k_smooth_force_fin, one, tb, 0, c->stream, c->B, 0);
            st = read_scalars(c);
            if (st) return st;
            const double fmax = h.sm_fmax;
            if (!std::isfinite(fmax)) {
                result = fail(c, PEARL_ERR_SMOOTHING, "nonfinite smoothing force");
                break;
            }
            if (fmax == 0) {
                rep->steps = step - 1;
                rep->converged = 1;
                rep->final_relative_energy_change = 0;
                rep->final_max_phase_change = 0;
                break;
            }
            double scale = std::min(base, max_phase_change / fmax);
            bool accepted = false;
            double e_new = previous, change = 0;
            for (int bt = 0; bt < 24 && !accepted; ++bt) {
                PEARL_LAUNCH(c, k_smooth_trial, grid, c->B, 0, scale);
                hipLaunchKernelGGL(k_smooth_trial_fin, one, tb, 0, c->stream, c->B, 0);
                st = read_scalars(c);
                if (st) return st;
                if (h.sm_bad > 0) {
                    result = fail(c, PEARL_ERR_PROJECTION, "Simplex projection failed");
                    break;
                }
                change = h.sm_change;
                e_new = h.sm_energy * P.dx * P.dy;
                const double etol = 1e-14 * std::max({1.0, std::abs(previous), std::abs(rep->energy_initial)});
                if (std::isfinite(e_new) && e_new <= previous + etol) accepted = true;
                else scale *= .5;
            }
            if (result) break;
            if (!accepted) {
                result = fail(c, PEARL_ERR_SMOOTHING, "sharp smoothing could not find an energy-decreasing step");
                break;
            }
            // the trial phases become current
            h.cur ^= 1;
            st = write_scalars(c);
            if (st) return st;
            const double rel = std::abs(previous - e_new) / std::max(std::abs(previous), 1e-300);
            rep->steps = step;
            rep->energy_final = e_new;
            rep->final_relative_energy_change = rel;
            rep->final_max_phase_change = change;
            previous = e_new;
            if (step >= min_steps && rel < energy_rtol && change < 1e-5) {
                rep->converged = 1;
                break;
            }
        }
        if (rep->steps == 0) rep->energy_final = rep->energy_initial;
        // carbon and mu stay where they were: follow the phases to the other buffer
        if (h.cur != cur0) {
            const size_t n = static_cast<size_t>(P.n);
            HIP_CHECK(c, hipMemcpyAsync(dom_c(c, 0, h.cur), dom_c(c, 0, cur0), n * sizeof(double), hipMemcpyDeviceToDevice,
                                        c->stream));
            HIP_CHECK(c, hipMemcpyAsync(dom_mu(c, 0, h.cur), dom_mu(c, 0, cur0), n * sizeof(double),
                                        hipMemcpyDeviceToDevice, c->stream));
        }
        h.e0_valid = 0;
        h.m0_valid = 0;
        st = write_scalars(c);
        if (st) return st;
        HIP_CHECK(c, hipStreamSynchronize(c->stream));
        return result;
    } catch (const std::bad_alloc&) {
        return alloc_fail(c);
    }
}

static_assert(sizeof(FillOp) == sizeof(pearl_fill) && offsetof(FillOp, p) == offsetof(pearl_fill, p) &&
                  offsetof(FillOp, dir) == offsetof(pearl_fill, dir),
              "FillOp mirrors pearl_fill");
static_assert(FILL_CUBE == PEARL_FILL_CUBE && FILL_CYLINDER == PEARL_FILL_CYLINDER &&
                  FILL_CENTERBOX == PEARL_FILL_CENTERBOX2D && FILL_YJUNCTION == PEARL_FILL_YJUNCTION2D &&
                  FILL_YJUNCTION_LAMELLAE == PEARL_FILL_YJUNCTIONLAMELLAE2D &&
                  FILL_ADHERED_ARM == PEARL_FILL_ADHEREDPEARLITEARM2D && FILL_ELLIPSE == PEARL_FILL_PEARLITEELLIPSE2D,
              "method codes of the C ABI");

int pearl_fill_phases(pearl_ctx* c, const pearl_fill* ops, int count) {
    try {
        int st = ctx_check(c, 0);
        if (st) return st;
        if (count < 0 || (count > 0 && !ops)) return fail(c, PEARL_ERR_ARGUMENT, "invalid FILL operation list");
        // What the kernel relies on: a known method, phase IDs 0..2 where a
        // label is taken from a parameter, finite parameters and directions, and
        // at least one pair on a pearlite arm.  The rest is read_config's.
        for (int k = 0; k < count; ++k) {
            const pearl_fill& f = ops[k];
            bool ok = f.method >= 0 && f.method < FILL_METHODS;
            for (double v : f.p) ok = ok && std::isfinite(v);
            for (const auto& u : f.dir) ok = ok && std::isfinite(u[0]) && std::isfinite(u[1]);
            auto phase = [&](int i) { ok = ok && (f.p[i] == 0.0 || f.p[i] == 1.0 || f.p[i] == 2.0); };
            if (ok) {
                switch (f.method) {
                    case FILL_CUBE:
                    case FILL_CYLINDER: phase(0); break;
                    case FILL_CENTERBOX: phase(4), phase(5); break;
                    case FILL_YJUNCTION: phase(11), phase(12); break;
                    case FILL_YJUNCTION_LAMELLAE: phase(13), phase(14), phase(15); break;
                    default: phase(9), phase(10); break;  // pearlite arm, ellipse
                }
                if (f.method == FILL_ADHERED_ARM) ok = ok && f.p[4] >= 1.0 && f.p[4] <= 2147483647.0;
            }
            if (!ok) return fail(c, PEARL_ERR_ARGUMENT, "invalid FILL operation " + std::to_string(k));
        }
        st = read_scalars(c);
        if (st) return st;
        const KP& P = c->kp[0];
        FillOp* d_ops = nullptr;
        if (count > 0) {
            HIP_CHECK(c, hipMalloc(&d_ops, static_cast<size_t>(count) * sizeof(FillOp)));
            const hipError_t e =
                hipMemcpyAsync(d_ops, ops, static_cast<size_t>(count) * sizeof(FillOp), hipMemcpyHostToDevice, c->stream);
            if (e != hipSuccess) {
                (void)hipFree(d_ops);
                return hip_fail(c, e, "hipMemcpyAsync(fill operations)");
            }
        }
        PEARL_LAUNCH(c, k_fill, dim3(static_cast<unsigned>(P.nb)), c->B, 0, d_ops, count);
        hipError_t e = hipGetLastError();
        const hipError_t e2 = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = e2;
        if (d_ops) (void)hipFree(d_ops);
        if (e != hipSuccess) return hip_fail(c, e, "k_fill");
        // the energy and mass of the previous phases no longer hold
        DevScal& h = c->hS[0];
        h.e0_valid = 0;
        h.m0_valid = 0;
        return write_scalars(c);
    } catch (const std::bad_alloc&) {
        return alloc_fail(c);
    }
}

int pearl_init_profile(pearl_ctx* c, int kind, const double* col, const double* row, const double* cell,
                       int phase_a, int phase_b) {
    try {
        int st = ctx_check(c, 0);
        if (st) return st;
        const KP& P = c->kp[0];
        const size_t nx = static_cast<size_t>(P.nx), ny = static_cast<size_t>(P.ny);
        bool ok = kind >= PEARL_PROFILE_LAMELLA && kind <= PEARL_PROFILE_CELLS;
        if (ok && kind != PEARL_PROFILE_LAMELLA)
            ok = phase_a >= 0 && phase_a <= 2 && phase_b >= 0 && phase_b <= 2 && phase_a != phase_b;
        // the arrays a kind reads: (pointer, length)
        const double* arr[2] = {nullptr, nullptr};
        size_t len[2] = {0, 0};
        if (kind == PEARL_PROFILE_LAMELLA) arr[0] = col, len[0] = nx, arr[1] = row, len[1] = ny;
        else if (kind == PEARL_PROFILE_FLAT) arr[0] = row, len[0] = ny;
        else arr[0] = cell, len[0] = nx * ny;
        for (int k = 0; ok && k < 2; ++k) {
            if (!len[k]) continue;
            if (!arr[k]) ok = false;
            for (size_t j = 0; ok && j < len[k]; ++j) ok = std::isfinite(arr[k][j]);
        }
        if (!ok) return fail(c, PEARL_ERR_ARGUMENT, "invalid profile");
        st = read_scalars(c);
        if (st) return st;
        double* d = nullptr;
        const size_t total = len[0] + len[1];
        HIP_CHECK(c, hipMalloc(&d, total * sizeof(double)));
        hipError_t e = hipMemcpyAsync(d, arr[0], len[0] * sizeof(double), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess && len[1])
            e = hipMemcpyAsync(d + len[0], arr[1], len[1] * sizeof(double), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) {
            const double* d_col = kind == PEARL_PROFILE_LAMELLA ? d : nullptr;
            const double* d_row = kind == PEARL_PROFILE_LAMELLA ? d + len[0] : (kind == PEARL_PROFILE_FLAT ? d : nullptr);
            const double* d_cell = kind == PEARL_PROFILE_CELLS ? d : nullptr;
            PEARL_LAUNCH(c, k_profile, dim3(static_cast<unsigned>(P.nb)), c->B, 0, kind, d_col, d_row, d_cell, phase_a,
                         phase_b);
            e = hipGetLastError();
        }
        const hipError_t e2 = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = e2;
        (void)hipFree(d);
        if (e != hipSuccess) return hip_fail(c, e, "k_profile");
        DevScal& h = c->hS[0];
        h.e0_valid = 0;
        h.m0_valid = 0;
        return write_scalars(c);
    } catch (const std::bad_alloc&) {
        return alloc_fail(c);
    }
}

int pearl_init_composition(pearl_ctx* c, int mode, double value, double* mu_out, double* carbon_out,
                           int* valid_out) {
    try {
        int st = ctx_check(c, 0);
        if (st) return st;
        if ((mode != PEARL_COMPOSITION_MU && mode != PEARL_COMPOSITION_MEAN) || !std::isfinite(value) || !mu_out ||
            !carbon_out || !valid_out)
            return fail(c, PEARL_ERR_ARGUMENT, "invalid composition set-up");
        st = read_scalars(c);
        if (st) return st;
        const KP& P = c->kp[0];
        const dim3 grid(static_cast<unsigned>(P.nb)), one(1), tb(TPB);
        DevScal& h = c->hS[0];
        double mu = value;
        if (mode == PEARL_COMPOSITION_MEAN) {  // :1814-1822
            PEARL_LAUNCH(c, k_init_sums, grid, c->B, 0);
            hipLaunchKernelGGL(k_init_sums_fin, one, tb, 0, c->stream, c->B, 0);
            st = read_scalars(c);
            if (st) return st;
            if (h.in_bad > 0) return fail(c, PEARL_ERR_MIXTURE, "Invalid mixture coefficients");
            mu = (static_cast<double>(P.nx * P.ny) * value - h.in_c0) / h.in_s;
        }
        PEARL_LAUNCH(c, k_init_carbon, grid, c->B, 0, mu);
        double valid = 0, mass = 0;
        st = state_direct(c, 0, true, &valid, &mass, nullptr);
        if (st) return st;
        *mu_out = mu;
        *carbon_out = mass;
        *valid_out = valid > 0.5 ? 1 : 0;
        h.e0_valid = 0;
        h.m0_valid = 0;
        return write_scalars(c);
    } catch (const std::bad_alloc&) {
        return alloc_fail(c);
    }
}

int pearl_reduced_energy(pearl_ctx* c, double* energy) {
    int st = ctx_check(c, 0);
    if (st) return st;
    if (!energy) return PEARL_ERR_ARGUMENT;
    return state_direct(c, 0, !c->prm[0].phase_only, nullptr, nullptr, energy);
}

int pearl_total_carbon(pearl_ctx* c, double* mass) {
    int st = ctx_check(c, 0);
    if (st) return st;
    if (!mass) return PEARL_ERR_ARGUMENT;
    return state_direct(c, 0, false, nullptr, mass, nullptr);
}

int pearl_valid_state(pearl_ctx* c, int* valid) {
    int st = ctx_check(c, 0);
    if (st) return st;
    if (!valid) return PEARL_ERR_ARGUMENT;
    double v = 0;
    st = state_direct(c, 0, false, &v, nullptr, nullptr);
    *valid = v > 0.5 ? 1 : 0;
    return st;
}

int pearl_get_layout(const pearl_ctx* c, int* tpb, int* cpt) {
    if (!c) return PEARL_ERR_ARGUMENT;
    if (tpb) *tpb = TPB;
    if (cpt) *cpt = c->k;
    return PEARL_OK;
}

int pearl_profile(pearl_ctx* c, int enable) {
    int st = ctx_check(c, 0);
    if (st) return st;
    st = read_scalars(c);
    if (st) return st;
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
    const int st = read_scalars(c);
    if (st) return st;
    if (launches) *launches = c->launches[kernel];
    if (total_ms) *total_ms = c->ms[kernel];
    if (bytes) *bytes = c->bytes[kernel];
    return PEARL_OK;
}

const char* pearl_kernel_name(int kernel) {
    if (kernel < 0 || kernel >= PEARL_KERNEL_COUNT) return "";
    return kKernelNames[kernel];
}

}  // extern "C"

namespace {
typedef double probe_vec __attribute__((ext_vector_type(2)));

template <bool NT>
__global__ __launch_bounds__(256) void k_stream_probe(const probe_vec* __restrict__ a, probe_vec* __restrict__ b,
                                                      long long n) {
    const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    if (NT) __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
    else b[i] = a[i];
}

// Milliseconds of `reps` copies of n vectors with kernel variant NT.
template <bool NT>
hipError_t time_copies(const probe_vec* a, probe_vec* b, long long n, int reps, hipStream_t st, hipEvent_t e0,
                       hipEvent_t e1, float* ms) {
    const dim3 grid(static_cast<unsigned>((n + 255) / 256));
    for (int r = 0; r < 3; ++r) hipLaunchKernelGGL(k_stream_probe<NT>, grid, dim3(256), 0, st, a, b, n);
    hipError_t e = hipEventRecord(e0, st);
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
        hipLaunchKernelGGL(k_stream_probe<NT>, grid, dim3(256), 0, st, a, b, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    return e;
}
}  // namespace

extern "C" {

int pearl_stream_probe(int device, long long bytes, int reps, double* gbps) {
    // one thread per vector: the grid's work-items must stay below 2^32
    if (!gbps || bytes <= 0 || bytes % 16 != 0 || bytes / 16 > (static_cast<long long>(1) << 31) || reps < 1)
        return PEARL_ERR_ARGUMENT;
    *gbps = 0;
    // The caller's current device is restored, and a failed call leaves no
    // error behind for the next hipGetLastError of another entry point.
    int prev = 0;
    if (hipGetDevice(&prev) != hipSuccess || hipSetDevice(device) != hipSuccess) {
        (void)hipGetLastError();
        return PEARL_ERR_DEVICE;
    }
    const long long n = bytes / 16;
    probe_vec *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&a, static_cast<size_t>(bytes));
    if (e == hipSuccess) e = hipMalloc(&b, static_cast<size_t>(bytes));
    if (e == hipSuccess) e = hipStreamCreate(&st);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemsetAsync(a, 0, static_cast<size_t>(bytes), st);
    float ms_plain = 0, ms_nt = 0;
    if (e == hipSuccess) e = time_copies<false>(a, b, n, reps, st, e0, e1, &ms_plain);
    if (e == hipSuccess) e = time_copies<true>(a, b, n, reps, st, e0, e1, &ms_nt);
    if (e == hipSuccess) {
        const double ms = std::min(ms_plain, ms_nt);
        *gbps = ms > 0 ? 2.0 * static_cast<double>(bytes) * reps / (ms * 1e-3) / 1e9 : 0.0;
    }
    if (e1) (void)hipEventDestroy(e1);
    if (e0) (void)hipEventDestroy(e0);
    if (st) (void)hipStreamDestroy(st);
    if (b) (void)hipFree(b);
    if (a) (void)hipFree(a);
    (void)hipSetDevice(prev);
    if (e != hipSuccess) (void)hipGetLastError();
    return e == hipSuccess ? PEARL_OK : (e == hipErrorOutOfMemory ? PEARL_ERR_ALLOC : PEARL_ERR_DEVICE);
}

}  // extern "C"

// The streams of a cell kernel without its arithmetic: `reads` planes read
// and `writes` planes written, nontemporal, one column of TY cells per
// thread over 256-column tiles, the planes spaced as in the arena (the cell
// kernels' geometry).  At most 8 of
// each; the counts are uniform, so their tests are scalar branches.
constexpr int kMixMax = 8;
template <int TY>
__global__ __launch_bounds__(TPB) void k_stream_mix(const double* const* __restrict__ in, double* const* __restrict__ out,
                                                  int reads, int writes, int nx, int ny, int tiles_x) {
    const int lb = static_cast<int>(blockIdx.x), tx = lb % tiles_x, ty = lb / tiles_x;
    const int x = tx * TPB + static_cast<int>(threadIdx.x), y0 = ty * TY;
    if (x >= nx) return;
#pragma unroll
    for (int r = 0; r < TY; ++r) {
        if (y0 + r >= ny) break;
        const long long i = static_cast<long long>(y0 + r) * nx + x;
        double v = 0;
#pragma unroll
        for (int q = 0; q < kMixMax; ++q)
            if (q < reads) v += __builtin_nontemporal_load(in[q] + i);
#pragma unroll
        for (int q = 0; q < kMixMax; ++q)
            if (q < writes) __builtin_nontemporal_store(v + q, out[q] + i);
    }
}

extern "C" int pearl_stream_mix(int device, int reads, int writes, int nx, int ny, int rows_per_thread, int reps,
                                double* us, double* gbps) {
    const bool rows_ok = rows_per_thread == 1 || rows_per_thread == 2 || rows_per_thread == 4 ||
                         rows_per_thread == 8 || rows_per_thread == 16;
    if (!us || !gbps || reads < 1 || reads > kMixMax || writes < 1 || writes > kMixMax || nx < 1 || ny < 1 ||
        !rows_ok || reps < 1 || static_cast<long long>(nx) * ny > (1LL << 31))
        return PEARL_ERR_ARGUMENT;
    *us = *gbps = 0;
    int prev = 0;
    if (hipGetDevice(&prev) != hipSuccess || hipSetDevice(device) != hipSuccess) {
        (void)hipGetLastError();
        return PEARL_ERR_DEVICE;
    }
    // one allocation with the planes spaced as in a context's arena (the
    // cells plus 576)
    const size_t n = static_cast<size_t>(nx) * static_cast<size_t>(ny), spacing = n + 576;
    std::vector<double*> planes(static_cast<size_t>(reads + writes), nullptr);
    double* mem = nullptr;
    double** d_ptr = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&mem, planes.size() * spacing * sizeof(double));
    for (size_t k = 0; k < planes.size() && e == hipSuccess; ++k) planes[k] = mem + k * spacing;
    if (e == hipSuccess) e = hipMalloc(&d_ptr, planes.size() * sizeof(double*));
    if (e == hipSuccess) e = hipStreamCreate(&st);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    for (int q = 0; q < reads && e == hipSuccess; ++q) e = hipMemsetAsync(planes[q], 0, n * sizeof(double), st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_ptr, planes.data(), planes.size() * sizeof(double*), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        const int tiles_x = (nx + TPB - 1) / TPB;
        const unsigned blocks = static_cast<unsigned>(tiles_x * ((ny + rows_per_thread - 1) / rows_per_thread));
        const double* const* in = d_ptr;
        double* const* out = d_ptr + reads;
        auto launch = [&] {
            switch (rows_per_thread) {
                case 1: hipLaunchKernelGGL(k_stream_mix<1>, dim3(blocks), dim3(TPB), 0, st, in, out, reads, writes, nx, ny, tiles_x); break;
                case 2: hipLaunchKernelGGL(k_stream_mix<2>, dim3(blocks), dim3(TPB), 0, st, in, out, reads, writes, nx, ny, tiles_x); break;
                case 4: hipLaunchKernelGGL(k_stream_mix<4>, dim3(blocks), dim3(TPB), 0, st, in, out, reads, writes, nx, ny, tiles_x); break;
                case 8: hipLaunchKernelGGL(k_stream_mix<8>, dim3(blocks), dim3(TPB), 0, st, in, out, reads, writes, nx, ny, tiles_x); break;
                default: hipLaunchKernelGGL(k_stream_mix<16>, dim3(blocks), dim3(TPB), 0, st, in, out, reads, writes, nx, ny, tiles_x); break;
            }
        };
        for (int r = 0; r < 3; ++r) launch();
        e = hipEventRecord(e0, st);
        for (int r = 0; r < reps; ++r) launch();
        if (e == hipSuccess) e = hipEventRecord(e1, st);
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e == hipSuccess && ms > 0) {
            *us = 1e3 * ms / reps;
            *gbps = static_cast<double>(reads + writes) * sizeof(double) * static_cast<double>(n) / (*us * 1e3);
        }
    }
    if (e1) (void)hipEventDestroy(e1);
    if (e0) (void)hipEventDestroy(e0);
    if (st) (void)hipStreamDestroy(st);
    if (d_ptr) (void)hipFree(d_ptr);
    if (mem) (void)hipFree(mem);
    (void)hipSetDevice(prev);
    if (e != hipSuccess) (void)hipGetLastError();
    return e == hipSuccess ? PEARL_OK : (e == hipErrorOutOfMemory ? PEARL_ERR_ALLOC : PEARL_ERR_DEVICE);
}

// ---------------------------------------------------------------------------
// pearl_diagnostics_noalloc: the snapshot reductions of a state held in the
// caller's device memory, for callers that cannot allocate device memory.
// Its device memory beyond the inputs -- the parameter block, the device
// scalars that receive the totals, and the tile partials of the blocked
// reduction -- is one file-scope array per device, `scratch`, held until the
// process ends.  Host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device (the array's name is a
// host-side stand-in, not a device address).  Every call on a device shares
// that device's array: calls that overlap on one device race.
// ---------------------------------------------------------------------------
namespace {
constexpr long long kNoAllocTiles = 65536;  // tiles of 256 x k cells
constexpr size_t kNoAllocKP = 0, kNoAllocS = 2048, kNoAllocPart = 4096;
constexpr size_t kNoAllocBytes = kNoAllocPart + 8 * static_cast<size_t>(kNoAllocTiles) * sizeof(double);
static_assert(sizeof(KP) <= kNoAllocS - kNoAllocKP && sizeof(DevScal) <= kNoAllocPart - kNoAllocS,
              "parameter block or device scalars outgrew their scratch slots");
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

int noalloc_fail(char* msg, int msg_len, int code, const char* text) {
    if (msg && msg_len > 0) std::snprintf(msg, static_cast<size_t>(msg_len), "%s", text);
    return code;
}

// The current device's scratch address, from hipGetSymbolAddress on its
// first use on that device.
hipError_t scratch_of_current_device(unsigned char** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<unsigned char*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}
}  // namespace

extern "C" {

int pearl_diagnostics_noalloc(const pearl_params* params, double mu_reference, int cells_per_thread,
                              const double* phi, const double* carbon, const double* mu, void* stream,
                              pearl_diag* out, char* msg, int msg_len) {
    if (msg && msg_len > 0) msg[0] = 0;
    if (!params || !phi || !carbon || !mu || !out) return noalloc_fail(msg, msg_len, PEARL_ERR_ARGUMENT, "null argument");
    if (!params_ok(params)) return noalloc_fail(msg, msg_len, PEARL_ERR_ARGUMENT, "invalid parameters");
    int k = cells_per_thread <= 0 ? 4 : cells_per_thread;
    if (k != 1 && k != 2 && k != 4 && k != 8 && k != 16)
        return noalloc_fail(msg, msg_len, PEARL_ERR_ARGUMENT, "cells_per_thread must be 1, 2, 4, 8 or 16");
    KP P = make_kp(params, mu_reference, k);
    if (P.nb > kNoAllocTiles) {
        char text[160];
        std::snprintf(text, sizeof(text), "domain of %d tiles exceeds the no-allocation scratch (%lld tiles of 256 x %d cells)",
                      P.nb, kNoAllocTiles, k);
        return noalloc_fail(msg, msg_len, PEARL_ERR_ARGUMENT, text);
    }
    unsigned char* base = nullptr;
    hipError_t e = scratch_of_current_device(&base);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return noalloc_fail(msg, msg_len, PEARL_ERR_DEVICE, hipGetErrorString(e));
    }
    // parameter block and cleared device scalars (cur = 0: the state below
    // is "current"); the host copies live until the stream is synchronised
    P.off = 0;
    P.poff = 0;
    P.boff = 0;
    DevScal hs;
    std::memset(&hs, 0, sizeof(hs));
    hs.phase = PH_IDLE;
    Bufs B{};
    for (int b = 0; b < 2; ++b) {
        B.phi[b] = const_cast<double*>(phi);
        B.c[b] = const_cast<double*>(carbon);
        B.mu[b] = const_cast<double*>(mu);
    }
    B.kp = reinterpret_cast<const KP*>(base + kNoAllocKP);
    B.S = reinterpret_cast<DevScal*>(base + kNoAllocS);
    B.part = reinterpret_cast<double*>(base + kNoAllocPart);
    B.ndom = 1;
    hipStream_t st = static_cast<hipStream_t>(stream);
    e = hipMemcpyAsync(base + kNoAllocKP, &P, sizeof(KP), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(base + kNoAllocS, &hs, sizeof(DevScal), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        const dim3 g(static_cast<unsigned>(P.nb)), tb(TPB);
        switch (k) {
            case 1: hipLaunchKernelGGL(k_diagnostics<1>, g, tb, 0, st, B, 0); break;
            case 2: hipLaunchKernelGGL(k_diagnostics<2>, g, tb, 0, st, B, 0); break;
            case 4: hipLaunchKernelGGL(k_diagnostics<4>, g, tb, 0, st, B, 0); break;
            case 8: hipLaunchKernelGGL(k_diagnostics<8>, g, tb, 0, st, B, 0); break;
            default: hipLaunchKernelGGL(k_diagnostics<16>, g, tb, 0, st, B, 0); break;
        }
        hipLaunchKernelGGL(k_diagnostics_fin, dim3(1), tb, 0, st, B, 0);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&hs, base + kNoAllocS, sizeof(DevScal), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return noalloc_fail(msg, msg_len, PEARL_ERR_DEVICE, hipGetErrorString(e));
    }
    if (!params->phase_only && (hs.err & ERRF_MIXTURE_STATE))
        return noalloc_fail(msg, msg_len, PEARL_ERR_MIXTURE, "Invalid mixture coefficients");
    out->c_min = hs.dmin_c;
    out->c_max = hs.dmax_c;
    out->mu_min = hs.dmin_mu;
    out->mu_max = hs.dmax_mu;
    out->phi_sum_error = hs.dsumerr;
    out->mass = hs.dmass * P.dx * P.dy;
    out->energy = hs.denergy * P.dx * P.dy;
    return PEARL_OK;
}

int pearl_counters(const pearl_ctx* c, long long* attempts, long long* rejected, long long* cg) {
    if (!c) return PEARL_ERR_ARGUMENT;
    long long a = 0, r = 0, g = 0;
    for (int d = 0; d < c->ndom; ++d) {
        a += c->hS[d].attempts;
        r += c->hS[d].rejected_attempts;
        g += c->hS[d].cg_all;
    }
    if (attempts) *attempts = a;
    if (rejected) *rejected = r;
    if (cg) *cg = g;
    return PEARL_OK;
}

}  // extern "C"
