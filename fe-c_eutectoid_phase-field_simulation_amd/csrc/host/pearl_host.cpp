// pearl_gpu: the drop-in program for the reference's pearl_serial.
//
// Same command line, input files, output directory and files as the
// reference program (pearl_serial.cpp:2184-2336): Input.in,
// thermodynamics.txt, initialization_report.txt, init_sharp / init_smoothed
// phase snapshots, frame_NNNNNN.h5 + .xmf at every output time,
// diagnostics.csv, checkpoint.h5 and COMPLETED.txt / STOPPED.txt; checkpoints
// restart in either program.  Every field operation runs on the GPU through
// the C ABI of libpearl_gpu.so (include/pearl_gpu.h):
//
//   initialization   pearl_fill_phases (FILL recipes, sharp_smooth boxes),
//                    pearl_smooth (capillarity-only pre-smoothing and the
//                    interface energy of diffuse profiles),
//                    pearl_init_composition (carbon and mu)
//   checkpoints      pearl_valid_state, pearl_total_carbon (restart checks)
//   step loop        pearl_run between output times, pearl_maybe_shift,
//                    pearl_diagnostics
//
// The host reads and writes files, owns the output cadence and evaluates
// the diffuse initial profiles, whose sine is the C library's (the same
// values as the reference's).  Modules: input_file.hpp (grammar),
// settings.hpp (keys, defaults, temperature laws, signature),
// fill_rules.hpp (FILL admission), frames.hpp (HDF5 / XMF).
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <csignal>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "frames.hpp"
#include "pearl_gpu.h"
#include "settings.hpp"

namespace pearl_host {
namespace {

volatile std::sig_atomic_t g_interrupted = 0;
void on_signal(int) { g_interrupted = 1; }

// ---------------------------------------------------------------------------
// Command line
// ---------------------------------------------------------------------------
struct Options {
    std::string input, output, restart;
    bool thermo_only = false;
    int px = 1, py = 1, device = 0;
    long long max_steps = -1;  // --max-steps: stop this run after that many accepted steps
};

Options parse_command_line(int argc, char** argv) {
    if (argc < 3)
        throw std::invalid_argument(
            "Usage: pearl input.in NEW_OUTPUT_DIR [--restart frame.h5] [--px N --py N] [--max-steps N] [--thermo-only]");
    Options o;
    o.input = argv[1];
    o.output = argv[2];
    // flags with a value, and what each sets
    const std::vector<std::pair<std::string, std::function<void(const std::string&)>>> valued = {
        {"--restart", [&](const std::string& v) { o.restart = v; }},
        {"--px", [&](const std::string& v) { o.px = std::stoi(v); }},
        {"--py", [&](const std::string& v) { o.py = std::stoi(v); }},
        {"--max-steps", [&](const std::string& v) { o.max_steps = std::stoll(v); }},
        {"--device", [&](const std::string& v) { o.device = std::stoi(v); }},  // no reference counterpart
    };
    for (int k = 3; k < argc; ++k) {
        const std::string flag = argv[k];
        if (flag == "--thermo-only") {
            o.thermo_only = true;
            continue;
        }
        // Initialization always runs on the device; the flag that once
        // selected it is still accepted.
        if (flag == "--gpu-init") continue;
        if (k + 1 == argc) throw std::invalid_argument("missing CLI value");
        const std::string value = argv[++k];
        auto it = std::find_if(valued.begin(), valued.end(), [&](const auto& f) { return f.first == flag; });
        if (it == valued.end()) throw std::invalid_argument("unknown CLI flag: " + flag);
        it->second(value);
    }
    if (o.thermo_only && !o.restart.empty()) throw std::invalid_argument("--thermo-only cannot be combined with --restart");
    return o;
}

// ---------------------------------------------------------------------------
// Device context
// ---------------------------------------------------------------------------
class Device {
public:
    Device(const Settings& cfg, double mu_reference, int device) {
        const pearl_params p = cfg.solver_params();
        check(pearl_create(&p, mu_reference, PEARL_FP64, device, 0, &ctx_));
    }
    ~Device() { pearl_destroy(ctx_); }
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;

    void check(int status) const {
        if (status != PEARL_OK) throw std::runtime_error(ctx_ ? pearl_last_error(ctx_) : "pearl_gpu call failed");
    }
    pearl_ctx* get() const { return ctx_; }

    void upload(const HostState& st) {
        check(pearl_set_state(ctx_, st.phi.data(), st.carbon.data(), st.mu.data(), &st.sc));
    }
    void download(HostState& st) {
        check(pearl_get_state(ctx_, st.phi.data(), st.carbon.data(), st.mu.data(), &st.sc));
    }
    void download_phases(HostState& st) { check(pearl_get_state(ctx_, st.phi.data(), nullptr, nullptr, nullptr)); }
    pearl_scalars scalars() {
        pearl_scalars sc{};
        check(pearl_get_state(ctx_, nullptr, nullptr, nullptr, &sc));
        return sc;
    }
    void set_scalars(const pearl_scalars& sc) { check(pearl_set_state(ctx_, nullptr, nullptr, nullptr, &sc)); }

private:
    pearl_ctx* ctx_ = nullptr;
};

// ---------------------------------------------------------------------------
// Reports
// ---------------------------------------------------------------------------
using Lines = std::vector<std::pair<std::string, std::string>>;

template <class T>
std::string num(T v) {
    std::ostringstream o;
    o << std::setprecision(17) << v;
    return o.str();
}

template <class... T>
std::string list(T... v) {
    std::ostringstream o;
    o << std::setprecision(17);
    const char* sep = "";
    ((o << sep << v, sep = ","), ...);
    return o.str();
}

void write_lines(const std::filesystem::path& path, const Lines& lines, const char* failure) {
    std::ofstream out(path);
    if (!out) throw std::runtime_error(std::string("cannot write ") + failure);
    for (const auto& kv : lines) out << kv.first << '=' << kv.second << '\n';
    out.close();
    if (!out) throw std::runtime_error(std::string(failure) + " write failed");
}

// thermodynamics.txt (write_thermodynamic_report :2118-2160).
Lines thermo_report(const Settings& c) {
    Lines L = {{"thermo_mode", c.thermo_mode},
               {"run_temperature_K", num(c.T)},
               {"Vm_m3_per_mol", num(c.Vm)},
               {"energy_convention", "legacy v0.3 molar-code convention; chemical term Psi/Vm, c=atomic fraction"},
               {"unit_scope",
                "this preserves the inherited conversion; it does not certify the material's physical molar volume"},
               {"thermal_history", "isothermal at the run temperature; no finite-rate cooling segment"}};
    const bool reference = c.thermo_mode == "eutectoid_reference";
    if (reference) {
        L.insert(L.end(), {{"T_eutectoid_K", num(c.T_eutectoid)},
                           {"undercooling_K", num(c.undercooling)},
                           {"max_undercooling_K", num(c.max_undercooling)},
                           {"temperature_law", c.temperature_law}});
        if (c.temperature_law == "phase_diagram_fixed_curvature") {
            const auto& s = c.diagram.slope;
            L.insert(L.end(),
                     {{"phase_diagram_slope_units", "carbon mole fraction/K (dc_eq/dT)"},
                      {"phase_diagram_curvature", "fixed at supplied reference Aeq"},
                      {"phase_diagram_gamma_reference", "ABC held at reference values"},
                      {"phase_diagram_reference_c_alpha_theta_gamma",
                       list(c.diagram.c_eq[0], c.diagram.c_eq[1], c.diagram.c_eq[2])},
                      {"phase_diagram_slopes_alpha_ag_gamma_ag_theta_tg_gamma_tg", list(s[0], s[1], s[2], s[3])},
                      {"phase_diagram_reference_residual", "preserved, not retuned"}});
        }
    }
    for (int p = 0; p < 3; ++p) {
        const std::string n = kPhaseNames[p];
        if (reference) {
            const Parabola& r = c.at_eutectoid[p];
            L.push_back({"reference_" + n + "_ABC", list(r.A, r.B, r.C)});
            if (c.temperature_law == "linear_gp") {
                const GrandPotential& d = c.dg_dT[p];
                L.push_back({"dq_dT_" + n + "_q2_q1_q0", list(d.q2, d.q1, d.q0)});
            }
        }
        const Parabola& f = c.thermo[p];
        const GrandPotential g = GrandPotential::of(f);
        L.push_back({"effective_" + n + "_ABC", list(f.A, f.B, f.C)});
        L.push_back({"effective_" + n + "_q2_q1_q0", list(g.q2, g.q1, g.q0)});
        L.push_back({"effective_" + n + "_chi", num(f.chi())});
        L.push_back({"D_" + n, num(f.D)});
    }
    L.push_back({"initial_composition_mode", c.composition_mode});
    L.push_back({"initial_mu_at_run_temperature", c.composition_mode != "mean"
                                                      ? num(c.initial_mu)
                                                      : "solved after smoothing from specified mean_carbon"});
    L.push_back({"initialization_note", "one local mu field; tip samples at distinct positions need not be equal"});
    L.push_back({"tau_at_ag_tg", list(c.tau[0], c.tau[1], c.tau[2])});
    L.push_back({"tau_note",
                 "explicit input, not recalibrated by changing temperature; physical applicability remains a separate "
                 "check"});
    return L;
}

// What the initialization reports (SmoothReport :843-852).
struct InitReport {
    int steps = 0;
    bool converged = false;
    double energy_initial = 0, energy_final = 0, relative_change = 0, max_phase_change = 0;
    std::array<double, 3> before{}, after{};
};

// Mean of each phase over the cells, summed cell by cell (phase_fractions :854-862).
std::array<double, 3> phase_fractions(const HostState& st) {
    std::array<double, 3> out{};
    const double n = static_cast<double>(st.nx) * st.ny;
    for (int a = 0; a < 3; ++a) {
        double s = 0;
        const double* v = st.plane(a);
        for (std::size_t i = 0; i < st.cells(); ++i) s += v[i];
        out[a] = s / n;
    }
    return out;
}

Lines init_report(const Settings& c, const HostState& st, const InitReport& r, double mu_reference) {
    return {{"initialization", c.initialization},
            {"physical_time_after_preprocess", num(st.sc.time)},
            {"solver_step_after_preprocess", num(st.sc.step)},
            {"smooth_steps_completed", num(r.steps)},
            {"smooth_energy_converged", r.converged ? "1" : "0"},
            {"smooth_energy_initial", num(r.energy_initial)},
            {"smooth_energy_final", num(r.energy_final)},
            {"smooth_final_relative_energy_change", num(r.relative_change)},
            {"smooth_final_max_phase_change", num(r.max_phase_change)},
            {"phase_fraction_before", list(r.before[0], r.before[1], r.before[2])},
            {"phase_fraction_after", list(r.after[0], r.after[1], r.after[2])},
            {"mu_reference", num(mu_reference)},
            {"NOTE",
             "smoothing is capillarity-only preprocessing pseudo-time; carbon diffusion is disabled and physical time is "
             "reset to zero before coupled evolution."}};
}

// ---------------------------------------------------------------------------
// Initial state
// ---------------------------------------------------------------------------

// The reference's diffuse interface profile across a signed distance d
// (profile :563-569): 1 below -half width, 0 above, a sine in between.
double profile(double d, double eps) {
    const double pi = std::acos(-1.0), half = pi * pi * eps / 8;
    return d <= -half ? 1.0 : (d >= half ? 0.0 : 0.5 * (1 - std::sin(4 * d / (pi * eps))));
}

// Phases of the diffuse initializations (:1766-1792) on the device.  The
// profile values take the C library's sine, so they are formed here and the
// device builds the phases from them (pearl_init_profile).  lamella and flat
// are separable: a profile per column (the cementite slab of theta_columns
// cells, periodic in x) and per row (the colony below seed_height), so each
// sine is evaluated once per column or row; circle needs one per cell.
void diffuse_phases(const Settings& c, const Device& dev) {
    const int nx = c.nx, ny = c.ny;
    auto centre_x = [&](int x) { return (x + .5) * c.dx; };
    auto centre_y = [&](int y) { return (y + .5) * c.dy; };
    std::vector<double> rows(ny);
    for (int y = 0; y < ny; ++y) rows[y] = profile(centre_y(y) - c.seed_height, c.eps);
    if (c.initialization == "lamella") {
        const double width = c.theta_columns * c.dx, length = nx * c.dx;
        std::vector<double> cols(nx);
        for (int x = 0; x < nx; ++x) {
            const double xx = centre_x(x);
            cols[x] = profile(xx < width ? -std::min(xx, width - xx) : std::min(xx - width, length - xx), c.eps);
        }
        dev.check(pearl_init_profile(dev.get(), PEARL_PROFILE_LAMELLA, cols.data(), rows.data(), nullptr, 0, 2));
    } else if (c.initialization == "flat") {
        dev.check(pearl_init_profile(dev.get(), PEARL_PROFILE_FLAT, nullptr, rows.data(), nullptr, c.phase_a,
                                     c.phase_b));
    } else {
        std::vector<double> cells(static_cast<std::size_t>(nx) * ny);
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x)
                cells[static_cast<std::size_t>(y) * nx + x] =
                    profile(std::hypot(centre_x(x) - .5 * nx * c.dx, centre_y(y) - .5 * ny * c.dy) - c.radius, c.eps);
        dev.check(pearl_init_profile(dev.get(), PEARL_PROFILE_CELLS, nullptr, nullptr, cells.data(), c.phase_a,
                                     c.phase_b));
    }
}

// ---------------------------------------------------------------------------
// The run
// ---------------------------------------------------------------------------
class Run {
public:
    Run(Options opt, Settings cfg) : opt_(std::move(opt)), cfg_(std::move(cfg)), out_(opt_.output) {}

    int execute() {
        create_output();
        if (opt_.thermo_only) {
            std::cout << "THERMO_ONLY: report written; no initialization or evolution performed\n";
            return 0;
        }
        mu_reference_ = opt_.restart.empty() ? initialize() : restore();
        if (state_.sc.time >= cfg_.end_time) throw std::runtime_error("checkpoint time >= end_time");
        if (!dev_) {
            dev_ = std::make_unique<Device>(cfg_, mu_reference_, opt_.device);
            dev_->upload(state_);
        }
        evolve();
        return 0;
    }

private:
    Options opt_;
    Settings cfg_;
    std::filesystem::path out_;
    HostState state_;
    double mu_reference_ = 0;
    std::unique_ptr<Device> dev_;
    // output cadence and loop statistics
    std::ofstream csv_;
    long long frames_ = 0, cg_total_ = 0;
    int cg_peak_ = 0;
    pearl_step_info last_{};
    std::chrono::steady_clock::time_point start_;

    void create_output() {
        bool ok = true;
        try {
            if (std::filesystem::exists(out_)) throw std::runtime_error("output path already exists; refuse overwrite");
            std::filesystem::create_directories(out_);
            std::filesystem::copy_file(opt_.input, out_ / "Input.in");
            write_lines(out_ / "thermodynamics.txt", thermo_report(cfg_), "thermodynamics report");
        } catch (...) {
            ok = false;
        }
        if (!ok) throw std::runtime_error("cannot create NEW output directory");
    }

    // initialize (:1761-1837) on a context of its own (mu_reference is what
    // the composition set-up returns), then the run's context receives the
    // state.
    double initialize() {
        state_.allocate(cfg_.nx, cfg_.ny);
        Device init(cfg_, cfg_.initial_mu, opt_.device);
        InitReport rep;
        const pearl_scalars zero{};
        if (cfg_.sharp_labels()) {
            const std::vector<pearl_fill> ops = sharp_label_ops(cfg_);
            init.check(pearl_fill_phases(init.get(), ops.data(), static_cast<int>(ops.size())));
            init.download_phases(state_);
            write_phases(out_ / "init_sharp.h5", state_, cfg_, "init_sharp", 0);
            rep.before = phase_fractions(state_);
            smooth(init, rep, cfg_.smooth_steps, cfg_.smooth_min_steps);
            init.download_phases(state_);
            rep.after = phase_fractions(state_);
            write_phases(out_ / "init_smoothed.h5", state_, cfg_, "init_smoothed", rep.steps);
        } else {
            diffuse_phases(cfg_, init);
            init.download_phases(state_);
            rep.before = rep.after = phase_fractions(state_);
            smooth(init, rep, 0, 0);  // the interface energy alone
            rep.converged = true;
        }
        const bool mean = cfg_.composition_mode == "mean";
        double mu = 0, carbon = 0;
        int valid = 0;
        init.check(pearl_init_composition(init.get(), mean ? PEARL_COMPOSITION_MEAN : PEARL_COMPOSITION_MU,
                                          mean ? cfg_.initial_carbon : cfg_.initial_mu, &mu, &carbon, &valid));
        init.check(pearl_get_state(init.get(), nullptr, state_.carbon.data(), state_.mu.data(), nullptr));
        state_.sc = zero;
        state_.sc.dt = cfg_.dt;
        state_.sc.initial_carbon = carbon;
        if (!valid) throw std::runtime_error("initial state out of model range");
        write_lines(out_ / "initialization_report.txt", init_report(cfg_, state_, rep, mu), "initialization report");
        return mu;
    }

    void smooth(Device& d, InitReport& rep, int steps, int min_steps) {
        pearl_smooth_report r{};
        d.check(pearl_smooth(d.get(), steps, min_steps, cfg_.smooth_energy_rtol, cfg_.smooth_max_phase_change, &r));
        rep.steps = r.steps;
        rep.converged = r.converged != 0;
        rep.energy_initial = r.energy_initial;
        rep.energy_final = r.energy_final;
        rep.relative_change = r.final_relative_energy_change;
        rep.max_phase_change = r.final_max_phase_change;
    }

    // A checkpoint of this run's physics (read_checkpoint :1970-2020): the
    // file's own checks, then the state's on the device (model range),
    // the closure c = composition(mu, mixture(phi)) and the carbon ledger.
    double restore() {
        const double mu_reference = read_frame(opt_.restart, cfg_, state_);
        dev_ = std::make_unique<Device>(cfg_, mu_reference, opt_.device);
        dev_->upload(state_);
        int valid = 0;
        dev_->check(pearl_valid_state(dev_->get(), &valid));
        if (!valid) throw std::runtime_error("invalid checkpoint state");
        if (closure_error() > cfg_.closure_tol) throw std::runtime_error("checkpoint closure failed");
        double carbon = 0;
        dev_->check(pearl_total_carbon(dev_->get(), &carbon));
        const pearl_scalars& s = state_.sc;
        const double ledger = carbon + s.carbon_removed - s.carbon_added - s.initial_carbon;
        if (!cfg_.phase_only() && std::abs(ledger) > 1e-10 * cfg_.nx * cfg_.ny * cfg_.dx * cfg_.dy)
            throw std::runtime_error("checkpoint mass ledger failed");
        return mu_reference;
    }

    // max |c - (chi mu + c0)| with chi, c0 of the interpolated parabolas at
    // each cell's phases (mixture :125-138, composition :139-141).
    double closure_error() const {
        const std::size_t n = state_.cells();
        double worst = 0;
        for (std::size_t i = 0; i < n; ++i) {
            const double p[3] = {state_.phi[i], state_.phi[n + i], state_.phi[2 * n + i]};
            const double pair3 = 2.0 * p[0] * p[1] * p[2];
            double chi = 0, c0 = 0, mobility = 0;
            for (int a = 0; a < 3; ++a) {
                const double h = p[a] * p[a] * (3.0 - 2.0 * p[a]) + pair3;
                const double x = cfg_.thermo[a].chi();
                chi += h * x;
                c0 -= h * x * cfg_.thermo[a].B;
                mobility += p[a] * cfg_.thermo[a].D * x;
            }
            if (!(chi > 0.0) || !std::isfinite(chi) || !std::isfinite(c0) || !std::isfinite(mobility) || mobility < 0.0)
                throw std::runtime_error("Invalid mixture coefficients");
            worst = std::max(worst, std::abs(state_.carbon[i] - std::fma(chi, state_.mu[i], c0)));
        }
        return worst;
    }

    // frame_NNNNNN.h5 and a diagnostics.csv row for the current state
    // (snapshot :2266-2292).
    void snapshot() {
        pearl_diag dg{};
        dev_->check(pearl_diagnostics(dev_->get(), &dg));
        dev_->download(state_);
        const pearl_scalars& s = state_.sc;
        const double ledger = dg.mass + s.carbon_removed - s.carbon_added - s.initial_carbon;
        const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - start_).count();
        char name[32];
        std::snprintf(name, sizeof(name), "frame_%06lld.h5", frames_++);
        write_frame(out_ / name, state_, cfg_, mu_reference_);
        const double row[] = {s.time,
                              static_cast<double>(s.step),
                              s.dt,
                              dg.mass / (cfg_.nx * cfg_.ny * cfg_.dx * cfg_.dy),
                              ledger,
                              dg.energy,
                              dg.phi_sum_error,
                              dg.c_min,
                              dg.c_max,
                              dg.mu_min,
                              dg.mu_max};
        csv_ << std::setprecision(17);
        for (double v : row) csv_ << v << ',';
        csv_ << last_.cg_iterations << ',' << cg_peak_ << ',' << last_.closure << ',' << s.rejected << ','
             << s.shift_distance << ',' << wall << '\n';
        csv_.flush();
        std::cout << "t=" << s.time << " step=" << s.step << " dt_next=" << s.dt << " mass_error=" << ledger
                  << " CG=" << last_.cg_iterations << " rejected=" << s.rejected << "\n";
    }

    // The step loop (:2249-2316).  Between two output times the steps run on
    // the device (pearl_run, at most kChunk per call so an interrupt is seen
    // between calls) and end where the reference's loop takes its output:
    // time >= next_output - eps.  At each output time: the clock snapped to
    // the output time within eps, the moving window, a snapshot.
    void evolve() {
        constexpr long long kChunk = 512;
        std::signal(SIGINT, on_signal);
        std::signal(SIGTERM, on_signal);
        start_ = std::chrono::steady_clock::now();
        csv_.open(out_ / "diagnostics.csv");
        if (!csv_) throw std::runtime_error("cannot open diagnostics");
        csv_ << "time,step,dt_next,mean_carbon,mass_ledger_error,reduced_energy,phi_sum_error,c_min,c_max,mu_min,mu_max,"
                "cg_last,cg_peak,closure_last,rejected,shift_distance,wall_s\n";
        std::cout << "resources: ranks=1 layout=1x1 grid=" << cfg_.nx << "x" << cfg_.ny << " device=" << opt_.device
                  << "\n";
        const double period = cfg_.output_dt;
        const double eps = 1e-12 * std::max(cfg_.end_time, period);
        const long long first_step = state_.sc.step;
        double next_output = (std::floor((state_.sc.time + 1e-12 * period) / period) + 1.0) * period;
        snapshot();
        double written_at = state_.sc.time;
        dev_->check(pearl_set_stop_tolerance(dev_->get(), eps));
        pearl_scalars sc = dev_->scalars();
        bool stopped = false;
        while (sc.time < cfg_.end_time) {
            const long long done = sc.step - first_step;
            if (g_interrupted || (opt_.max_steps > 0 && done >= opt_.max_steps)) {
                stopped = true;
                break;
            }
            if (sc.step >= cfg_.max_steps) throw std::runtime_error("configured max_steps reached before end_time");
            long long budget = std::min<long long>(kChunk, cfg_.max_steps - sc.step);
            if (opt_.max_steps > 0) budget = std::min(budget, opt_.max_steps - done);
            long long taken = 0;
            dev_->check(pearl_run(dev_->get(), std::min(cfg_.end_time, next_output), budget, &taken, &cg_total_,
                                  &cg_peak_, &last_));
            sc = dev_->scalars();
            if (sc.time < next_output - eps && sc.time < cfg_.end_time - eps) continue;
            // snap to the output time, then (from there) to the end time
            double snapped = sc.time;
            if (std::abs(snapped - next_output) <= eps) snapped = next_output;
            if (std::abs(snapped - cfg_.end_time) <= eps) snapped = cfg_.end_time;
            if (snapped != sc.time) {
                sc.time = snapped;
                dev_->set_scalars(sc);
            }
            int shifted = 0;
            double front = -1;
            dev_->check(pearl_maybe_shift(dev_->get(), cfg_.moving_window(), cfg_.shift_trigger, cfg_.shift_target,
                                          cfg_.shift_gamma_carbon, &shifted, &front));
            snapshot();
            sc = state_.sc;
            written_at = sc.time;
            next_output += period;
        }
        if (sc.time != written_at) snapshot();
        dev_->download(state_);
        write_frame(out_ / "checkpoint.h5", state_, cfg_, mu_reference_);
        std::ofstream fin(out_ / (stopped ? "STOPPED.txt" : "COMPLETED.txt"));
        fin << std::setprecision(17) << "status=" << (stopped ? "USER_STOPPED" : "HORIZON_COMPLETED")
            << "\ntime=" << state_.sc.time << "\nstep=" << state_.sc.step << "\nCG_total=" << cg_total_
            << "\nrejected=" << state_.sc.rejected << "\nThis is a solver status, not certification of cooperative growth.\n";
    }
};

int program(int argc, char** argv) {
    const Options opt = parse_command_line(argc, argv);
    Settings cfg = read_settings(opt.input);
    if (opt.max_steps == 0 || opt.max_steps < -1) throw std::invalid_argument("max-steps must be positive");
    if (cfg.nx < 3 || cfg.ny < 2 || cfg.dx <= 0 || cfg.dy <= 0) throw std::invalid_argument("invalid grid");
    // the device context's size limit (no reference counterpart)
    if (3 * static_cast<long long>(cfg.nx) * cfg.ny >= (1LL << 31) - (1LL << 20))
        throw std::invalid_argument("grid too large for one device context (3*nx*ny must stay below 2^31 - 2^20)");
    if (opt.px != 1 || opt.py != 1) throw std::invalid_argument("serial build requires px=py=1");
    return Run(opt, std::move(cfg)).execute();
}

}  // namespace
}  // namespace pearl_host

int main(int argc, char** argv) {
    try {
        return pearl_host::program(argc, argv);
    } catch (const std::exception& e) {
        std::cerr << "ERROR: " << e.what() << '\n';
        return 1;
    }
}
