// Test-only harness around the reference solver (pearl_serial.cpp), compiled
// from the reference source where it lies (the include path is supplied by
// oracle/Makefile).  It exposes a C ABI so that the tests can drive the
// reference's own read_config / initialize / advance / maybe_shift functions
// on caller-supplied states and compare them against the oracle restatement
// and the HIP path.  Nothing in the product links or loads this library.
//
// The reference's main() is renamed at compile time (-Dmain=...), so this
// harness is a plain shared library.  Reference functions used:
//   read_config        pearl_serial.cpp:1600
//   initialize         pearl_serial.cpp:1761 (also up to its init_sharp observer call)
//   propose_phi        pearl_serial.cpp:961
//   implicit_diffusion pearl_serial.cpp:1010
//   valid_state        pearl_serial.cpp:1096
//   total_carbon       pearl_serial.cpp:784
//   reduced_energy     pearl_serial.cpp:793
//   advance            pearl_serial.cpp:1119
//   maybe_shift        pearl_serial.cpp:2053
//   read_checkpoint / write_frame  pearl_serial.cpp:1970 / :1914
#include "pearl_serial.cpp"

#include <cstring>
#include <memory>

namespace {
struct RefHandle {
    pearl::Config cfg;
    std::unique_ptr<pearl::Domain> dom;
    std::unique_ptr<pearl::State> st;
    double mu_reference = 0;
    std::string err;
};

void set_err(RefHandle* h, const std::exception& e) { h->err = e.what(); }

// Planar [3][ny][nx] <-> the reference's ghosted interleaved Array.
void put_phi(const pearl::Domain& d, pearl::Array& a, const double* p3) {
    const std::size_t n = static_cast<std::size_t>(d.nx) * d.ny;
    for (int y = 1; y <= d.ny; ++y)
        for (int x = 1; x <= d.nx; ++x)
            for (int c = 0; c < 3; ++c)
                a(x, y, c) = p3[c * n + static_cast<std::size_t>(y - 1) * d.nx + (x - 1)];
    d.halo(a);
}
void get_phi(const pearl::Domain& d, const pearl::Array& a, double* p3) {
    const std::size_t n = static_cast<std::size_t>(d.nx) * d.ny;
    for (int y = 1; y <= d.ny; ++y)
        for (int x = 1; x <= d.nx; ++x)
            for (int c = 0; c < 3; ++c)
                p3[c * n + static_cast<std::size_t>(y - 1) * d.nx + (x - 1)] = a(x, y, c);
}
void put_scalar(const pearl::Domain& d, pearl::Array& a, const double* v) {
    for (int y = 1; y <= d.ny; ++y)
        for (int x = 1; x <= d.nx; ++x)
            a(x, y) = v[static_cast<std::size_t>(y - 1) * d.nx + (x - 1)];
    d.halo(a);
}
void get_scalar(const pearl::Domain& d, const pearl::Array& a, double* v) {
    for (int y = 1; y <= d.ny; ++y)
        for (int x = 1; x <= d.nx; ++x)
            v[static_cast<std::size_t>(y - 1) * d.nx + (x - 1)] = a(x, y);
}
}  // namespace

extern "C" {

// Parse an input file with the reference parser.  Returns NULL on error and
// copies the message into err.
void* ref_open(const char* path, char* err, int errlen) {
    auto h = std::make_unique<RefHandle>();
    try {
        h->cfg = pearl::read_config(path);
        h->dom = std::make_unique<pearl::Domain>(h->cfg.nx, h->cfg.ny, h->cfg.dx, h->cfg.dy, 1, 1);
        h->st = std::make_unique<pearl::State>(*h->dom);
        h->st->dt = h->cfg.dt;
    } catch (const std::exception& e) {
        if (err && errlen > 0) {
            std::strncpy(err, e.what(), errlen - 1);
            err[errlen - 1] = 0;
        }
        return nullptr;
    }
    return h.release();
}

void ref_close(void* hp) { delete static_cast<RefHandle*>(hp); }

const char* ref_last_error(void* hp) { return static_cast<RefHandle*>(hp)->err.c_str(); }

// Parsed parameters in a fixed layout (see oracle/oracle.h, ORC_PARAM_*):
//  [0..11]  A,B,C,D for alpha, theta, gamma
//  [12] eps [13] sigma [14] triple [15..17] tau_at, tau_ag, tau_tg [18] Vm
//  [19] dx [20] dy [21] nx [22] ny
//  [23] dt_max [24] dt_min [25] max_phase_change [26] cg_rtol [27] cg_atol
//  [28] closure_tol [29] energy_rtol [30] energy_atol [31] cg_max
//  [32] max_retries [33] check_energy [34] phase_only
//  [35] dt [36] end_time [37] output_dt [38] moving_window [39] shift_trigger
//  [40] shift_target [41] shift_gamma_carbon [42] max_steps [43] initial_mu
//  [44] T [45] initial_carbon
int ref_params(void* hp, double* out) {
    auto* h = static_cast<RefHandle*>(hp);
    const auto& c = h->cfg;
    for (int p = 0; p < 3; ++p) {
        out[4 * p + 0] = c.thermo[p].free_energy_quadratic;
        out[4 * p + 1] = c.thermo[p].free_energy_linear;
        out[4 * p + 2] = c.thermo[p].free_energy_constant;
        out[4 * p + 3] = c.thermo[p].diffusivity;
    }
    out[12] = c.phase.interface.epsilon;
    out[13] = c.phase.interface.sigma;
    out[14] = c.phase.interface.triple;
    for (int i = 0; i < 3; ++i) out[15 + i] = c.phase.tau[i];
    out[18] = c.phase.Vm;
    out[19] = c.dx;
    out[20] = c.dy;
    out[21] = c.nx;
    out[22] = c.ny;
    out[23] = c.control.dt_max;
    out[24] = c.control.dt_min;
    out[25] = c.control.max_phase_change;
    out[26] = c.control.cg_rtol;
    out[27] = c.control.cg_atol;
    out[28] = c.control.closure_tol;
    out[29] = c.control.energy_rtol;
    out[30] = c.control.energy_atol;
    out[31] = c.control.cg_max;
    out[32] = c.control.max_retries;
    out[33] = c.control.check_energy ? 1 : 0;
    out[34] = c.control.phase_only ? 1 : 0;
    out[35] = c.dt;
    out[36] = c.end_time;
    out[37] = c.output_dt;
    out[38] = c.moving_window ? 1 : 0;
    out[39] = c.shift_trigger;
    out[40] = c.shift_target;
    out[41] = c.shift_gamma_carbon;
    out[42] = static_cast<double>(c.max_steps);
    out[43] = c.initial_mu;
    out[44] = c.T;
    out[45] = c.initial_carbon;
    return 46;
}

int ref_signature(void* hp, char* buf, int len) {
    auto* h = static_cast<RefHandle*>(hp);
    const std::string s = h->cfg.physics_signature();
    if (static_cast<int>(s.size()) + 1 > len) return -static_cast<int>(s.size() + 1);
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return static_cast<int>(s.size());
}

// Full reference initialization (filling, capillarity smoothing, carbon).
int ref_initialize(void* hp) {
    auto* h = static_cast<RefHandle*>(hp);
    try {
        h->mu_reference = pearl::initialize(*h->dom, *h->st, h->cfg);
    } catch (const std::exception& e) {
        set_err(h, e);
        return -1;
    }
    return 0;
}

// Sharp labels of the reference's initialize (:1764-1797): the phases its
// observer receives at "init_sharp", before the smoothing.  The observer
// leaves initialize by throwing a private type, so the smoothing and the
// carbon set-up do not run (the handle's state is then partly initialized;
// reopen before other use).  Returns -2 for an initialization without
// sharp labels (lamella, flat, circle).
namespace {
struct SharpCaptured {};
}  // namespace
int ref_sharp_labels(void* hp, double* phi3) {
    auto* h = static_cast<RefHandle*>(hp);
    bool got = false;
    try {
        pearl::initialize(*h->dom, *h->st, h->cfg, nullptr,
                          [&](const pearl::Domain& d, const pearl::State& s, const char* stage, int) {
                              if (std::string(stage) != "init_sharp") return;
                              get_phi(d, s.phi, phi3);
                              got = true;
                              throw SharpCaptured{};
                          });
    } catch (const SharpCaptured&) {
    } catch (const std::exception& e) {
        set_err(h, e);
        return -1;
    }
    return got ? 0 : -2;
}

// Scalars: [0] time [1] dt [2] shift_distance [3] initial_carbon
// [4] carbon_added [5] carbon_removed [6] step [7] shifts [8] rejected
// [9] mu_reference
int ref_get_state(void* hp, double* phi3, double* carbon, double* mu, double* sc) {
    auto* h = static_cast<RefHandle*>(hp);
    const auto& d = *h->dom;
    auto& s = *h->st;
    if (phi3) get_phi(d, s.phi, phi3);
    if (carbon) get_scalar(d, s.carbon, carbon);
    if (mu) get_scalar(d, s.mu, mu);
    if (sc) {
        sc[0] = s.time; sc[1] = s.dt; sc[2] = s.shift_distance; sc[3] = s.initial_carbon;
        sc[4] = s.carbon_added; sc[5] = s.carbon_removed;
        sc[6] = static_cast<double>(s.step); sc[7] = static_cast<double>(s.shifts);
        sc[8] = static_cast<double>(s.rejected); sc[9] = h->mu_reference;
    }
    return 0;
}

int ref_set_state(void* hp, const double* phi3, const double* carbon, const double* mu, const double* sc) {
    auto* h = static_cast<RefHandle*>(hp);
    const auto& d = *h->dom;
    auto& s = *h->st;
    if (phi3) put_phi(d, s.phi, phi3);
    if (carbon) put_scalar(d, s.carbon, carbon);
    if (mu) put_scalar(d, s.mu, mu);
    if (sc) {
        s.time = sc[0]; s.dt = sc[1]; s.shift_distance = sc[2]; s.initial_carbon = sc[3];
        s.carbon_added = sc[4]; s.carbon_removed = sc[5];
        s.step = static_cast<long long>(sc[6]); s.shifts = static_cast<long long>(sc[7]);
        s.rejected = static_cast<long long>(sc[8]); h->mu_reference = sc[9];
    }
    return 0;
}

// Override numerical controls after parsing (tests force rejection paths).
int ref_set_controls(void* hp, const double* c10) {
    auto* h = static_cast<RefHandle*>(hp);
    auto& k = h->cfg.control;
    k.dt_max = c10[0]; k.dt_min = c10[1]; k.max_phase_change = c10[2];
    k.cg_rtol = c10[3]; k.cg_atol = c10[4]; k.closure_tol = c10[5];
    k.energy_rtol = c10[6]; k.energy_atol = c10[7];
    k.cg_max = static_cast<int>(c10[8]); k.max_retries = static_cast<int>(c10[9]);
    return 0;
}

// One reference advance.  info: accepted_dt, max_phase_change,
// linear_residual, closure, energy_change, cg_iterations, retries.
int ref_advance(void* hp, double stop_time, double* info) {
    auto* h = static_cast<RefHandle*>(hp);
    try {
        auto r = pearl::advance(*h->dom, *h->st, h->cfg.thermo, h->cfg.phase, h->cfg.control,
                                h->mu_reference, stop_time);
        info[0] = r.accepted_dt; info[1] = r.max_phase_change; info[2] = r.linear_residual;
        info[3] = r.closure; info[4] = r.energy_change; info[5] = r.cg_iterations;
        info[6] = r.retries;
    } catch (const std::exception& e) {
        set_err(h, e);
        return -1;
    }
    return 0;
}

// The reference main loop (pearl_serial.cpp:2296-2316) without file output:
// advance to min(end_time, next_output), snap the clock, apply the moving
// window at output times.  Stops after max_steps accepted steps or at
// end_time.  stats: [0] steps taken [1] cg_total [2] cg_peak [3] shifts done
int ref_run(void* hp, double end_time, long long max_steps, double* stats) {
    auto* h = static_cast<RefHandle*>(hp);
    try {
        auto& s = *h->st;
        const auto& cfg = h->cfg;
        double next_output = (std::floor((s.time + 1e-12 * cfg.output_dt) / cfg.output_dt) + 1.0) * cfg.output_dt;
        long long taken = 0, cg_total = 0, shifts = 0;
        int cg_peak = 0;
        while (s.time < end_time && taken < max_steps) {
            auto last = pearl::advance(*h->dom, s, cfg.thermo, cfg.phase, cfg.control, h->mu_reference,
                                       std::min(end_time, next_output));
            ++taken;
            cg_total += last.cg_iterations;
            cg_peak = std::max(cg_peak, last.cg_iterations);
            double eps = 1e-12 * std::max(end_time, cfg.output_dt);
            if (s.time >= next_output - eps || s.time >= end_time - eps) {
                if (std::abs(s.time - next_output) <= eps) s.time = next_output;
                if (std::abs(s.time - end_time) <= eps) s.time = end_time;
                if (pearl::maybe_shift(*h->dom, s, cfg)) ++shifts;
                next_output += cfg.output_dt;
            }
        }
        stats[0] = static_cast<double>(taken);
        stats[1] = static_cast<double>(cg_total);
        stats[2] = cg_peak;
        stats[3] = static_cast<double>(shifts);
    } catch (const std::exception& e) {
        set_err(h, e);
        return -1;
    }
    return 0;
}

double ref_propose_phi(void* hp, double dt, double* trial_phi3) {
    auto* h = static_cast<RefHandle*>(hp);
    pearl::State trial = *h->st;
    double ch = pearl::propose_phi(*h->dom, *h->st, trial, h->cfg.thermo, h->cfg.phase, dt);
    get_phi(*h->dom, trial.phi, trial_phi3);
    return ch;
}

// Implicit diffusion from the current state with a caller-supplied trial phi.
// info: ok, it, residual, closure.
int ref_implicit_diffusion(void* hp, const double* trial_phi3, double dt, double* tc, double* tmu, double* info) {
    auto* h = static_cast<RefHandle*>(hp);
    try {
        pearl::State trial = *h->st;
        put_phi(*h->dom, trial.phi, trial_phi3);
        auto di = pearl::implicit_diffusion(*h->dom, *h->st, trial, h->cfg.thermo, dt, h->mu_reference, h->cfg.control);
        info[0] = di.ok ? 1 : 0; info[1] = di.it; info[2] = di.residual; info[3] = di.closure;
        get_scalar(*h->dom, trial.carbon, tc);
        get_scalar(*h->dom, trial.mu, tmu);
    } catch (const std::exception& e) {
        set_err(h, e);
        return -1;
    }
    return 0;
}

double ref_reduced_energy(void* hp) {
    auto* h = static_cast<RefHandle*>(hp);
    return pearl::reduced_energy(*h->dom, *h->st, h->cfg.thermo, h->cfg.phase, h->mu_reference,
                                 h->cfg.control.phase_only);
}

double ref_total_carbon(void* hp) {
    auto* h = static_cast<RefHandle*>(hp);
    return pearl::total_carbon(*h->dom, *h->st);
}

int ref_valid_state(void* hp) {
    auto* h = static_cast<RefHandle*>(hp);
    return pearl::valid_state(*h->dom, *h->st, h->cfg.thermo, h->cfg.control.phase_only) ? 1 : 0;
}

// Moving window with explicit trigger/target (the parsed ones otherwise).
int ref_maybe_shift(void* hp, int* shifted) {
    auto* h = static_cast<RefHandle*>(hp);
    try {
        *shifted = pearl::maybe_shift(*h->dom, *h->st, h->cfg) ? 1 : 0;
    } catch (const std::exception& e) {
        set_err(h, e);
        return -1;
    }
    return 0;
}

int ref_set_window(void* hp, int on, double trigger, double target, double c_gamma) {
    auto* h = static_cast<RefHandle*>(hp);
    h->cfg.moving_window = on != 0;
    h->cfg.shift_trigger = trigger;
    h->cfg.shift_target = target;
    h->cfg.shift_gamma_carbon = c_gamma;
    return 0;
}

// Smoothing alone (used to pin the preprocessing restatement).
int ref_smooth(void* hp, int max_steps, int min_steps, double rtol, double max_change, double* report) {
    auto* h = static_cast<RefHandle*>(hp);
    try {
        auto r = pearl::smooth_interface_only(*h->dom, *h->st, h->cfg.phase, max_steps, min_steps, rtol, max_change);
        report[0] = r.steps; report[1] = r.converged; report[2] = r.energy_initial;
        report[3] = r.energy_final; report[4] = r.final_relative_energy_change;
        report[5] = r.final_max_phase_change;
    } catch (const std::exception& e) {
        set_err(h, e);
        return -1;
    }
    return 0;
}

// Write a checkpoint with the reference writer (restart-compatibility tests).
int ref_write_frame(void* hp, const char* path) {
    auto* h = static_cast<RefHandle*>(hp);
    try {
        pearl::write_frame(*h->dom, *h->st, h->cfg, h->mu_reference, path);
    } catch (const std::exception& e) {
        set_err(h, e);
        return -1;
    }
    return 0;
}

int ref_read_checkpoint(void* hp, const char* path) {
    auto* h = static_cast<RefHandle*>(hp);
    try {
        h->mu_reference = pearl::read_checkpoint(*h->dom, *h->st, h->cfg, path);
    } catch (const std::exception& e) {
        set_err(h, e);
        return -1;
    }
    return 0;
}

}  // extern "C"
