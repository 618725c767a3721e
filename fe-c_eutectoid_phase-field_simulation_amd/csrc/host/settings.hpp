// Run settings of the drop-in program: what the reference's read_config
// (pearl_serial.cpp:1544-1758) accepts, read through a key schema.
//
// The schema is a list of sections; a section is an ordered list of items,
// each either a key (its kind, where it is stored, and its default: none,
// a constant, or a value computed from the keys read before it) or a rule
// over what has been read so far.  One interpreter walks the sections, so
// parsing, defaults, the order of the checks and the record of which keys
// were read come from the same table.  Error texts, defaults and the order in
// which keys are examined are the reference's (a file with several problems
// reports the same first one); the temperature laws keep the reference's
// long-double expressions (:186-276), which fix the bits of the effective
// thermodynamic coefficients.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <functional>
#include <iomanip>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "fill_rules.hpp"
#include "input_file.hpp"
#include "pearl_gpu.h"

namespace pearl_host {

// The checkpoint signature names the inherited numerical model by this
// digest (pearl_serial.cpp:31-33); checkpoints are interchangeable with the
// reference program's only when the signatures are equal, so the value is
// the same.
constexpr const char* kModelDigest = "344eef3b2e1fdbd8a41d81378491b923f6d815d8ef917f8397729c5a57965580";
constexpr const char* kPhaseNames[3] = {"alpha", "theta", "gamma"};

// Free energy of one phase, f(c) = A c^2 + B c + C, and its diffusivity.
struct Parabola {
    double A = 0, B = 0, C = 0, D = 0;
    void require_valid() const {
        if (!std::isfinite(A) || !std::isfinite(B) || !std::isfinite(C) || !std::isfinite(D) || A <= 0.0 || D < 0.0)
            throw std::invalid_argument("Parabola requires finite A>0, B, C, D>=0");
    }
    double mu_at(double c) const { return std::fma(2.0 * A, c, B); }  // df/dc
    double c_at(double mu) const { return (mu - B) / (2.0 * A); }
    double chi() const { return 1.0 / (2.0 * A); }
};
using Thermo = std::array<Parabola, 3>;

// The grand potential of a parabola, g(mu) = min_c f(c) - mu c =
// q2 mu^2 + q1 mu + q0 with q2 = -1/(4A), q1 = B/(2A), q0 = C - B^2/(4A), in
// long double.  The linear temperature law moves g's coefficients linearly
// in T and maps them back.
struct GrandPotential {
    long double q2 = 0, q1 = 0, q0 = 0;

    static GrandPotential of(const Parabola& f) {
        f.require_valid();
        const long double A = f.A, B = f.B, C = f.C;
        return {-1 / (4 * A), B / (2 * A), C - B * B / (4 * A)};
    }
    Parabola to_parabola(double D) const {
        if (!std::isfinite(q2) || !std::isfinite(q1) || !std::isfinite(q0) || q2 >= 0)
            throw std::invalid_argument("temperature law requires finite q2<0, q1, q0");
        const long double A = -1 / (4 * q2), B = -q1 / (2 * q2);
        Parabola f{static_cast<double>(A), static_cast<double>(B), static_cast<double>(q0 + B * B / (4 * A)), D};
        f.require_valid();
        return f;
    }
};

inline void require_undercooling(double undercooling) {
    if (!std::isfinite(undercooling) || undercooling < 0)
        throw std::invalid_argument("undercooling must be finite and nonnegative");
}

// linear_gp: g(T) = g(T_e) + (T - T_e) dg/dT, T - T_e = -undercooling.
inline Parabola linear_gp_law(const Parabola& at_eutectoid, const GrandPotential& dg_dT, double undercooling) {
    require_undercooling(undercooling);
    if (!std::isfinite(dg_dT.q2) || !std::isfinite(dg_dT.q1) || !std::isfinite(dg_dT.q0))
        throw std::invalid_argument("nonfinite temperature derivative");
    at_eutectoid.require_valid();
    if (undercooling == 0) return at_eutectoid;
    GrandPotential g = GrandPotential::of(at_eutectoid);
    const long double dT = -static_cast<long double>(undercooling);
    g.q2 += dT * dg_dT.q2;
    g.q1 += dT * dg_dT.q1;
    g.q0 += dT * dg_dT.q0;
    return g.to_parabola(at_eutectoid.D);
}

// phase_diagram_fixed_curvature: the curvatures A stay at their eutectoid
// values; the ferrite and cementite parabolas (B, C) move so that their
// common tangents with the austenite parabola touch at the contact
// compositions extrapolated along the phase-diagram slopes dc/dT (for
// product p: c_p on the p/gamma boundary and c_gamma on it).  The austenite
// parabola is the reference.
struct PhaseDiagram {
    std::array<double, 3> c_eq{};     // alpha, theta, gamma contacts at T_e
    std::array<double, 4> slope{};    // alpha on a/g, gamma on a/g, theta on t/g, gamma on t/g
};

inline Thermo phase_diagram_law(const Thermo& at_eutectoid, const PhaseDiagram& pd, double undercooling) {
    require_undercooling(undercooling);
    for (const Parabola& f : at_eutectoid) f.require_valid();
    for (double c : pd.c_eq)
        if (!std::isfinite(c) || c < 0 || c > 1) throw std::invalid_argument("phase-diagram reference carbon must be in [0,1]");
    for (double s : pd.slope)
        if (!std::isfinite(s)) throw std::invalid_argument("phase-diagram slope must be finite dc/dT");
    const long double dT = -static_cast<long double>(undercooling);
    const long double A_g = at_eutectoid[2].A, c_g = pd.c_eq[2];
    Thermo out = at_eutectoid;
    for (int p = 0; p < 2; ++p) {
        const long double A_p = at_eutectoid[p].A, c_p = pd.c_eq[p];
        const long double shift_p = dT * pd.slope[2 * p], shift_g = dT * pd.slope[2 * p + 1];
        const long double moved_p = c_p + shift_p, moved_g = c_g + shift_g;
        if (!std::isfinite(moved_p) || !std::isfinite(moved_g) || moved_p < 0 || moved_p > 1 || moved_g < 0 ||
            moved_g > 1)
            throw std::invalid_argument("phase-diagram extrapolated carbon outside [0,1]");
        if (undercooling == 0) continue;
        out[p].B = static_cast<double>(at_eutectoid[p].B + 2 * (A_g * shift_g - A_p * shift_p));
        out[p].C = static_cast<double>(at_eutectoid[p].C + A_p * shift_p * (2 * c_p + shift_p) -
                                       A_g * shift_g * (2 * c_g + shift_g));
        out[p].require_valid();
    }
    return out;
}

struct Settings {
    // grid and interface
    int nx = 0, ny = 0;
    double dx = 0, dy = 0, eps = 0, sigma = 0, triple = 0, Vm = 0;
    std::array<double, 3> tau{};
    // thermodynamics
    std::string thermo_mode, temperature_law = "linear_gp";
    double T = 0, T_eutectoid = 0, undercooling = 0, max_undercooling = 0;
    Thermo thermo{}, at_eutectoid{};
    std::array<GrandPotential, 3> dg_dT{};
    PhaseDiagram diagram;
    // time and numerical controls
    double dt = 0, dt_max = 0, dt_min = 0, end_time = 0, output_dt = 0;
    int max_steps = 100000000;
    double max_phase_change = .04, cg_rtol = 1e-11, cg_atol = 1e-14, closure_tol = 1e-10, energy_atol = 0;
    int cg_max = 2000;
    // initialization
    std::string initialization = "lamella", mode = "coupled";
    double seed_height = 0, radius = 0;
    int theta_columns = 0, phase_a = 0, phase_b = 2;
    int smooth_steps = 100, smooth_min_steps = 100;
    double smooth_energy_rtol = 1e-10, smooth_max_phase_change = 0.02;
    std::vector<FillLine> fill_lines;
    std::vector<pearl_fill> fills;  // the admitted FILL operations (microsim)
    // composition
    std::string composition_mode;  // "mu", "gamma" or "mean"
    double initial_mu = 0, initial_carbon = 0;
    // moving window
    int moving_window_flag = 0;
    double shift_trigger = 0, shift_target = 0, shift_gamma_carbon = 0;

    bool moving_window() const { return moving_window_flag != 0; }
    bool phase_only() const { return mode == "phase_only"; }
    bool sharp_labels() const { return initialization == "sharp_smooth" || initialization == "microsim"; }

    // The C ABI's parameter block.
    pearl_params solver_params() const {
        pearl_params k{};
        k.nx = nx;
        k.ny = ny;
        k.dx = dx;
        k.dy = dy;
        for (int p = 0; p < 3; ++p) {
            k.A[p] = thermo[p].A;
            k.B[p] = thermo[p].B;
            k.C[p] = thermo[p].C;
            k.D[p] = thermo[p].D;
            k.tau[p] = tau[p];
        }
        k.eps = eps;
        k.sigma = sigma;
        k.triple = triple;
        k.Vm = Vm;
        k.dt_max = dt_max;
        k.dt_min = dt_min;
        k.max_phase_change = max_phase_change;
        k.cg_rtol = cg_rtol;
        k.cg_atol = cg_atol;
        k.closure_tol = closure_tol;
        k.energy_rtol = 1e-11;  // not an input key (:771)
        k.energy_atol = energy_atol;
        k.cg_max = cg_max;
        k.max_retries = 24;
        k.check_energy = 1;
        k.phase_only = phase_only() ? 1 : 0;
        return k;
    }

    // physics_signature (:1566-1597): every value that defines the physics,
    // ';'-separated, 17 significant digits.  Stored in every frame and
    // compared on restart.
    std::string signature() const {
        std::ostringstream o;
        o << std::setprecision(17);
        auto put = [&](const auto& v) { o << v << ';'; };
        o << "pearlite-v0.4-reference;source=" << kModelDigest << ";3ph;binary;equal-sigma;Jat=0;Xperiodic;Ynoflux;";
        for (double v : {static_cast<double>(nx), static_cast<double>(ny), dx, dy, Vm, eps, sigma, triple}) put(v);
        put(thermo_mode);
        for (double v : {T, T_eutectoid, undercooling, max_undercooling}) put(v);
        if (thermo_mode == "eutectoid_reference") {
            for (const Parabola& f : at_eutectoid)
                for (double v : {f.A, f.B, f.C}) put(v);
            if (temperature_law == "linear_gp") {
                for (const GrandPotential& s : dg_dT)
                    for (long double v : {s.q2, s.q1, s.q0}) put(v);
            } else {
                put("phase-diagram-fixed-curvature-v1");
                for (double c : diagram.c_eq) put(c);
                for (double s : diagram.slope) put(s);
            }
        }
        for (const Parabola& f : thermo)
            for (double v : {f.A, f.B, f.C, f.D}) put(v);
        for (double t : tau) put(t);
        put(mode);
        put(initialization);
        put(seed_height);
        put(theta_columns);
        put(radius);
        put(phase_a);
        put(phase_b);
        put(composition_mode);
        put(initial_mu);
        put(initial_carbon);
        put(smooth_steps);
        put(smooth_min_steps);
        put(smooth_energy_rtol);
        put(smooth_max_phase_change);
        put(moving_window());
        put(shift_trigger);
        put(shift_target);
        o << shift_gamma_carbon;
        if (initialization == "microsim") {
            o << ";filling-v1";
            for (const FillLine& f : fill_lines) {
                o << ';' << f.method;
                for (double v : f.p) o << ';' << v;
            }
        }
        return o.str();
    }
};

// ---------------------------------------------------------------------------
// The key schema and its interpreter
// ---------------------------------------------------------------------------
namespace schema {

using Fallback = std::function<double(const Settings&)>;
using Rule = std::function<void(Settings&, InputFile&)>;

struct Item {
    enum Kind { REAL, INTEGER, TEXT, RULE } kind = RULE;
    std::string key;
    double Settings::*real = nullptr;
    int Settings::*integer = nullptr;
    std::string Settings::*text = nullptr;
    Fallback fallback;           // none: the key is required (REAL, INTEGER)
    std::string text_default;    // TEXT
    Rule rule;                   // RULE
};

inline Item real(std::string key, double Settings::*m, Fallback f = nullptr) {
    Item i;
    i.kind = Item::REAL;
    i.key = std::move(key);
    i.real = m;
    i.fallback = std::move(f);
    return i;
}
inline Item integer(std::string key, int Settings::*m, Fallback f = nullptr) {
    Item i;
    i.kind = Item::INTEGER;
    i.key = std::move(key);
    i.integer = m;
    i.fallback = std::move(f);
    return i;
}
inline Item text(std::string key, std::string Settings::*m, std::string def) {
    Item i;
    i.kind = Item::TEXT;
    i.key = std::move(key);
    i.text = m;
    i.text_default = std::move(def);
    return i;
}
inline Item rule(Rule r) {
    Item i;
    i.kind = Item::RULE;
    i.rule = std::move(r);
    return i;
}
inline Fallback constant(double v) {
    return [v](const Settings&) { return v; };
}

// A number as the reference reads one: std::stod's grammar and errors, the
// whole text, finite.  An absent key takes its default; a key without a
// default, or whose default is not finite, is required.
inline double read_real(InputFile& in, const std::string& key, const Settings& s, const Fallback& fallback) {
    const std::string* t = in.take(key);
    if (!t) {
        const double def = fallback ? fallback(s) : std::numeric_limits<double>::quiet_NaN();
        if (!std::isfinite(def)) throw std::runtime_error("missing required key: " + key);
        return def;
    }
    std::size_t used = 0;
    const double v = std::stod(*t, &used);
    if (used != t->size() || !std::isfinite(v)) throw std::runtime_error("invalid number: " + key);
    return v;
}

// An integer-valued number of at most 1e8 in magnitude; a negative default
// makes the key required.
inline int read_integer(InputFile& in, const std::string& key, const Settings& s, const Fallback& fallback) {
    Fallback f;
    if (fallback) {
        const double def = fallback(s);
        if (def >= 0) f = constant(def);
    }
    const double v = read_real(in, key, s, f);
    if (v != std::floor(v) || std::abs(v) > 1e8) throw std::runtime_error("invalid integer: " + key);
    return static_cast<int>(v);
}

inline void run(const std::vector<Item>& items, Settings& s, InputFile& in) {
    for (const Item& it : items) {
        switch (it.kind) {
            case Item::REAL: s.*it.real = read_real(in, it.key, s, it.fallback); break;
            case Item::INTEGER: s.*it.integer = read_integer(in, it.key, s, it.fallback); break;
            case Item::TEXT: {
                const std::string* t = in.take(it.key);
                s.*it.text = t ? *t : it.text_default;
                break;
            }
            case Item::RULE: it.rule(s, in); break;
        }
    }
}

// A rule that throws `message` unless `ok` holds.
template <class E = std::runtime_error>
inline Item require(std::function<bool(const Settings&)> ok, std::string message) {
    return rule([ok = std::move(ok), message = std::move(message)](Settings& s, InputFile&) {
        if (!ok(s)) throw E(message);
    });
}

inline double grid_length_max(const Settings& s) { return std::max(s.nx * s.dx, s.ny * s.dy); }

// Thermodynamics at the run temperature: eutectoid reference plus a
// temperature law, or coefficients given at the temperature.
inline void read_thermo(Settings& s, InputFile& in) {
    const std::string* mode = in.take("thermo_mode");
    s.thermo_mode = mode ? *mode : "";
    auto num = [&](const std::string& key) { return read_real(in, key, s, nullptr); };
    if (s.thermo_mode == "eutectoid_reference") {
        s.T_eutectoid = num("T_eutectoid");
        s.undercooling = num("undercooling");
        s.max_undercooling = num("max_undercooling");
        const std::string* law = in.take("temperature_law");
        s.temperature_law = law ? *law : "";
        if (s.temperature_law != "linear_gp" && s.temperature_law != "phase_diagram_fixed_curvature")
            throw std::runtime_error("temperature_law must be linear_gp or phase_diagram_fixed_curvature");
        s.T = s.T_eutectoid - s.undercooling;
        if (s.T_eutectoid <= 0 || s.T <= 0 || s.undercooling < 0 || s.max_undercooling < 0 ||
            s.undercooling > s.max_undercooling)
            throw std::runtime_error("temperature outside declared undercooling interval");
        const bool linear = s.temperature_law == "linear_gp";
        for (int p = 0; p < 3; ++p) {
            const std::string n = kPhaseNames[p];
            s.at_eutectoid[p] = {num("Aeq_" + n), num("Beq_" + n), num("Ceq_" + n), num("D_" + n)};
            if (!linear) continue;
            s.dg_dT[p] = {num("dq2_dT_" + n), num("dq1_dT_" + n), num("dq0_dT_" + n)};
            s.thermo[p] = linear_gp_law(s.at_eutectoid[p], s.dg_dT[p], s.undercooling);
        }
        if (!linear) {
            s.diagram.c_eq = {num("ceq_alpha"), num("ceq_theta"), num("ceq_gamma")};
            s.diagram.slope = {num("slope_alpha_on_alpha_gamma"), num("slope_gamma_on_alpha_gamma"),
                               num("slope_theta_on_theta_gamma"), num("slope_gamma_on_theta_gamma")};
            s.thermo = phase_diagram_law(s.at_eutectoid, s.diagram, s.undercooling);
        }
    } else if (s.thermo_mode == "direct_at_temperature") {
        s.T = num("temperature");
        if (s.T <= 0) throw std::runtime_error("temperature must be positive");
        for (int p = 0; p < 3; ++p) {
            const std::string n = kPhaseNames[p];
            s.thermo[p] = {num("A_" + n), num("B_" + n), num("C_" + n), num("D_" + n)};
            s.thermo[p].require_valid();
        }
    } else {
        throw std::runtime_error(
            "Set thermo_mode=eutectoid_reference or direct_at_temperature. "
            "Legacy inputs require an explicit mode and temperature; no temperature is guessed.");
    }
}

// Exactly one composition key; mu of the gamma composition from the
// austenite parabola.
inline void read_composition(Settings& s, InputFile& in) {
    const int given = static_cast<int>(in.has("initial_mu")) + static_cast<int>(in.has("gamma_carbon")) +
                      static_cast<int>(in.has("mean_carbon"));
    if (given != 1) throw std::runtime_error("specify exactly one of initial_mu, gamma_carbon, mean_carbon");
    if (in.has("initial_mu")) {
        s.composition_mode = "mu";
        s.initial_mu = read_real(in, "initial_mu", s, nullptr);
    } else if (in.has("gamma_carbon")) {
        s.composition_mode = "gamma";
        s.initial_carbon = read_real(in, "gamma_carbon", s, nullptr);
        s.initial_mu = s.thermo[2].mu_at(s.initial_carbon);
    } else {
        s.composition_mode = "mean";
        s.initial_carbon = read_real(in, "mean_carbon", s, nullptr);
    }
}

inline const std::vector<Item>& sections() {
    static const std::vector<Item> items = [] {
        std::vector<Item> v = {
            // grid and interface
            integer("nx", &Settings::nx), integer("ny", &Settings::ny), real("dx", &Settings::dx),
            real("dy", &Settings::dy, [](const Settings& s) { return s.dx; }), real("epsilon", &Settings::eps),
            real("sigma", &Settings::sigma), real("triple", &Settings::triple), real("Vm", &Settings::Vm),
            rule([](Settings& s, InputFile& in) {
                for (int p = 0; p < 3; ++p)
                    s.tau[p] = read_real(in, p == 0 ? "tau_at" : (p == 1 ? "tau_ag" : "tau_tg"), s, nullptr);
            }),
            // thermodynamics
            rule(read_thermo),
            // time
            real("dt", &Settings::dt), real("dt_max", &Settings::dt_max, [](const Settings& s) { return s.dt; }),
            real("dt_min", &Settings::dt_min, [](const Settings& s) { return s.dt / 1048576.0; }),
            real("end_time", &Settings::end_time), real("output_dt", &Settings::output_dt),
            integer("max_steps", &Settings::max_steps, constant(100000000)),
            // initialization
            text("initialization", &Settings::initialization, "lamella"), text("mode", &Settings::mode, "coupled"),
            require([](const Settings& s) { return s.mode == "coupled" || s.mode == "phase_only"; },
                    "mode must be coupled or phase_only"),
            real("seed_height", &Settings::seed_height, [](const Settings& s) { return s.ny * s.dy / 4; }),
            integer("theta_columns", &Settings::theta_columns, [](const Settings& s) { return s.nx / 2; }),
            real("radius", &Settings::radius, [](const Settings& s) { return std::min(s.nx * s.dx, s.ny * s.dy) / 4; }),
            integer("phase_a", &Settings::phase_a, constant(0)), integer("phase_b", &Settings::phase_b, constant(2)),
            require(
                [](const Settings& s) {
                    for (const char* m : {"lamella", "sharp_smooth", "flat", "circle", "microsim"})
                        if (s.initialization == m) return true;
                    return false;
                },
                "unknown initialization"),
            rule([](Settings& s, InputFile&) {
                if (s.initialization == "microsim") {
                    if (s.fill_lines.empty())
                        throw std::runtime_error("initialization=microsim requires at least one FILL operation");
                    s.fills = admit_fills(s.fill_lines, s.nx, s.ny);
                } else if (!s.fill_lines.empty()) {
                    throw std::runtime_error(
                        "FILL operations require initialization=microsim; no geometry is silently ignored");
                }
            }),
            integer("smooth_steps", &Settings::smooth_steps, constant(100)),
            integer("smooth_min_steps", &Settings::smooth_min_steps,
                    [](const Settings& s) { return static_cast<double>(s.smooth_steps); }),
            real("smooth_energy_rtol", &Settings::smooth_energy_rtol, constant(1e-10)),
            real("smooth_max_phase_change", &Settings::smooth_max_phase_change, constant(0.02)),
            require(
                [](const Settings& s) {
                    return s.phase_a >= 0 && s.phase_a <= 2 && s.phase_b >= 0 && s.phase_b <= 2 && s.phase_a != s.phase_b;
                },
                "invalid phase pair"),
            // composition
            rule(read_composition),
            // numerical controls
            real("max_phase_change", &Settings::max_phase_change, constant(.04)),
            real("cg_rtol", &Settings::cg_rtol, constant(1e-11)), real("cg_atol", &Settings::cg_atol, constant(1e-14)),
            integer("cg_max", &Settings::cg_max, constant(2000)),
            real("closure_tol", &Settings::closure_tol, constant(1e-10)),
            real("energy_atol", &Settings::energy_atol,
                 [](const Settings& s) { return 1e-13 * s.sigma * grid_length_max(s); }),
            // moving window
            integer("moving_window", &Settings::moving_window_flag, constant(0)),
            real("shift_trigger", &Settings::shift_trigger, [](const Settings& s) { return .65 * s.ny * s.dy; }),
            real("shift_target", &Settings::shift_target, [](const Settings& s) { return .35 * s.ny * s.dy; }),
            real("shift_gamma_carbon", &Settings::shift_gamma_carbon,
                 [](const Settings& s) {
                     return s.composition_mode == "gamma" ? s.initial_carbon : s.thermo[2].c_at(s.initial_mu);
                 }),
            rule([](Settings&, InputFile& in) { in.reject_unread(); }),
            // ranges
            require<std::invalid_argument>(
                [](const Settings& s) {
                    return s.eps > 0 && s.sigma > 0 && s.triple >= 0 && std::isfinite(s.eps) &&
                           std::isfinite(s.sigma) && std::isfinite(s.triple);
                },
                "Invalid interface parameters"),
            require<std::invalid_argument>(
                [](const Settings& s) {
                    for (double t : s.tau)
                        if (!(t > 0) || !std::isfinite(t)) return false;
                    return true;
                },
                "tau must be finite and positive"),
            require<std::invalid_argument>([](const Settings& s) { return s.Vm > 0 && std::isfinite(s.Vm); },
                                           "Invalid Vm"),
            require(
                [](const Settings& s) {
                    return s.nx >= 3 && s.ny >= 2 && s.dx > 0 && s.dy > 0 && s.dt > 0 && s.dt_min > 0 &&
                           s.dt_max >= s.dt && s.dt_min <= s.dt && s.end_time > 0 && s.output_dt > 0 && s.max_steps >= 1;
                },
                "invalid geometry/time bounds"),
            require(
                [](const Settings& s) {
                    return s.theta_columns >= 1 && s.theta_columns < s.nx && s.seed_height > 0 &&
                           s.seed_height < s.ny * s.dy;
                },
                "invalid seed geometry"),
            require(
                [](const Settings& s) {
                    return s.max_phase_change > 0 && s.max_phase_change <= .2 && s.cg_rtol > 0 && s.cg_atol > 0 &&
                           s.closure_tol > 0 && s.energy_atol >= 0;
                },
                "invalid numerical control"),
            require(
                [](const Settings& s) {
                    return s.smooth_steps >= 1 && s.smooth_min_steps >= 0 && s.smooth_min_steps <= s.smooth_steps &&
                           s.smooth_energy_rtol > 0 && s.smooth_max_phase_change > 0 && s.smooth_max_phase_change <= .2;
                },
                "invalid smoothing control"),
            require(
                [](const Settings& s) {
                    return !s.moving_window() ||
                           (s.shift_target > 0 && s.shift_trigger > s.shift_target && s.shift_trigger < s.ny * s.dy &&
                            s.shift_gamma_carbon >= 0 && s.shift_gamma_carbon <= 1);
                },
                "invalid moving window settings"),
        };
        return v;
    }();
    return items;
}

}  // namespace schema

// The sharp labels of an input as pearl_fill operations: the admitted FILL
// lines of initialization = microsim (:1772-1776), or sharp_smooth's two
// FILLCUBE boxes, theta in the first theta_columns columns and alpha in the
// rest, over the rows whose centre lies below seed_height (:1777-1786).
inline std::vector<pearl_fill> sharp_label_ops(const Settings& c) {
    if (c.initialization == "microsim") return c.fills;
    if (c.initialization != "sharp_smooth")
        throw std::invalid_argument("initialization = " + c.initialization + " has no sharp labels");
    int rows = 0;
    while (rows < c.ny && (rows + .5) * c.dy < c.seed_height) ++rows;
    std::vector<pearl_fill> boxes(2);
    const double spans[2][3] = {{1, 0, static_cast<double>(c.theta_columns - 1)},
                                {0, static_cast<double>(c.theta_columns), static_cast<double>(c.nx - 1)}};
    for (int b = 0; b < 2; ++b) {
        boxes[b] = pearl_fill{};
        boxes[b].method = PEARL_FILL_CUBE;
        boxes[b].p[0] = spans[b][0];
        boxes[b].p[1] = spans[b][1];
        boxes[b].p[4] = spans[b][2];
        boxes[b].p[5] = rows - 1;
    }
    return boxes;
}

inline Settings read_settings(const std::string& path) {
    InputFile in = InputFile::read(path);
    Settings s;
    s.fill_lines = in.fills();
    schema::run(schema::sections(), s, in);
    return s;
}

}  // namespace pearl_host
