// libpearl_input.so: the drop-in program's input reader (settings.hpp)
// behind the C ABI of include/pearl_input.h, so that every host reads input
// files through this one implementation.
#include "pearl_input.h"

#include <cstdio>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "settings.hpp"

namespace {

using pearl_host::Settings;

int fail(char* err, int err_len, const char* what) {
    if (err && err_len > 0) std::snprintf(err, static_cast<std::size_t>(err_len), "%s", what);
    return -1;
}

// A JSON object written field by field; reals as exact hexadecimal strings.
class Object {
public:
    Object& real(const char* key, double v) {
        char b[48];
        std::snprintf(b, sizeof b, "\"%a\"", v);
        return raw(key, b);
    }
    Object& integer(const char* key, long long v) { return raw(key, std::to_string(v)); }
    Object& text(const char* key, const std::string& v) { return raw(key, quoted(v)); }
    Object& reals(const char* key, const double* v, std::size_t n) { return raw(key, list(v, n)); }
    Object& raw(const char* key, const std::string& value) {
        body_ += body_.empty() ? "{" : ",";
        body_ += quoted(key) + ":" + value;
        return *this;
    }
    std::string str() const { return body_.empty() ? "{}" : body_ + "}"; }

    static std::string list(const double* v, std::size_t n) {
        std::string l = "[";
        for (std::size_t i = 0; i < n; ++i) {
            char b[48];
            std::snprintf(b, sizeof b, "%s\"%a\"", i ? "," : "", v[i]);
            l += b;
        }
        return l + "]";
    }

    static std::string quoted(const std::string& s) {
        std::string q = "\"";
        for (const char ch : s) {
            const unsigned char u = static_cast<unsigned char>(ch);
            if (ch == '"' || ch == '\\') {
                q += '\\';
                q += ch;
            } else if (u < 0x20) {
                char b[8];
                std::snprintf(b, sizeof b, "\\u%04x", u);
                q += b;
            } else {
                q += ch;
            }
        }
        return q + "\"";
    }

private:
    std::string body_;
};

// [[A, B, C, D] x 3]
std::string parabolas(const pearl_host::Thermo& t) {
    std::string l = "[";
    for (int p = 0; p < 3; ++p) {
        const double v[4] = {t[p].A, t[p].B, t[p].C, t[p].D};
        l += (p ? "," : "") + Object::list(v, 4);
    }
    return l + "]";
}

std::string to_json(const Settings& s) {
    Object o;
    o.integer("nx", s.nx).integer("ny", s.ny).real("dx", s.dx).real("dy", s.dy);
    o.real("epsilon", s.eps).real("sigma", s.sigma).real("triple", s.triple).real("Vm", s.Vm);
    o.reals("tau", s.tau.data(), 3);
    o.text("thermo_mode", s.thermo_mode).text("temperature_law", s.temperature_law);
    o.real("temperature", s.T).real("T_eutectoid", s.T_eutectoid).real("undercooling", s.undercooling);
    o.real("max_undercooling", s.max_undercooling);
    o.raw("thermo", parabolas(s.thermo));
    if (s.thermo_mode == "eutectoid_reference") {
        o.raw("at_eutectoid", parabolas(s.at_eutectoid));
        if (s.temperature_law == "linear_gp") {
            // read as doubles, held as long double: the conversion back is exact
            std::string l = "[";
            for (int p = 0; p < 3; ++p) {
                const double v[3] = {static_cast<double>(s.dg_dT[p].q2), static_cast<double>(s.dg_dT[p].q1),
                                     static_cast<double>(s.dg_dT[p].q0)};
                l += (p ? "," : "") + Object::list(v, 3);
            }
            o.raw("dg_dT", l + "]");
        } else {
            o.reals("c_eq", s.diagram.c_eq.data(), 3).reals("slopes", s.diagram.slope.data(), 4);
        }
    }
    o.real("dt", s.dt).real("dt_max", s.dt_max).real("dt_min", s.dt_min).real("end_time", s.end_time);
    o.real("output_dt", s.output_dt).integer("max_steps", s.max_steps);
    o.text("initialization", s.initialization).text("mode", s.mode);
    o.real("seed_height", s.seed_height).integer("theta_columns", s.theta_columns).real("radius", s.radius);
    o.integer("phase_a", s.phase_a).integer("phase_b", s.phase_b);
    o.integer("smooth_steps", s.smooth_steps).integer("smooth_min_steps", s.smooth_min_steps);
    o.real("smooth_energy_rtol", s.smooth_energy_rtol).real("smooth_max_phase_change", s.smooth_max_phase_change);
    std::string fills = "[";
    for (std::size_t k = 0; k < s.fill_lines.size(); ++k) {
        Object f;
        f.text("method", s.fill_lines[k].method).reals("p", s.fill_lines[k].p.data(), s.fill_lines[k].p.size());
        fills += (k ? "," : "") + f.str();
    }
    o.raw("fill", fills + "]");
    o.text("composition_mode", s.composition_mode).real("initial_mu", s.initial_mu);
    o.real("initial_carbon", s.initial_carbon);
    o.real("max_phase_change", s.max_phase_change).real("cg_rtol", s.cg_rtol).real("cg_atol", s.cg_atol);
    o.integer("cg_max", s.cg_max).real("closure_tol", s.closure_tol).real("energy_atol", s.energy_atol);
    o.integer("moving_window", s.moving_window_flag).real("shift_trigger", s.shift_trigger);
    o.real("shift_target", s.shift_target).real("shift_gamma_carbon", s.shift_gamma_carbon);
    o.text("physics_signature", s.signature());
    return o.str();
}

pearl_host::Parabola parabola(const double* v) { return {v[0], v[1], v[2], v[3]}; }

}  // namespace

extern "C" long long pearl_read_input(const char* path, char* json, long long json_len, char* err, int err_len) {
    if (!path) return fail(err, err_len, "no input path");
    try {
        const std::string text = to_json(pearl_host::read_settings(path));
        const long long n = static_cast<long long>(text.size());
        if (json && json_len > n) std::memcpy(json, text.c_str(), text.size() + 1);
        return n;
    } catch (const std::exception& e) {
        return fail(err, err_len, e.what());
    }
}

extern "C" int pearl_input_fills(const char* path, pearl_fill* ops, int max_ops, int* count, char* err,
                                 int err_len) {
    if (!path || !count || max_ops < 0 || (max_ops > 0 && !ops)) return fail(err, err_len, "invalid argument");
    try {
        const std::vector<pearl_fill> v = pearl_host::sharp_label_ops(pearl_host::read_settings(path));
        *count = static_cast<int>(v.size());
        for (int k = 0; k < *count && k < max_ops; ++k) ops[k] = v[static_cast<std::size_t>(k)];
        return 0;
    } catch (const std::exception& e) {
        return fail(err, err_len, e.what());
    }
}

extern "C" int pearl_linear_gp_law(const double at_eutectoid[4], const double dg_dT[3], double undercooling,
                                   double out[4], char* err, int err_len) {
    if (!at_eutectoid || !dg_dT || !out) return fail(err, err_len, "invalid argument");
    try {
        const pearl_host::GrandPotential g{dg_dT[0], dg_dT[1], dg_dT[2]};
        const pearl_host::Parabola f = pearl_host::linear_gp_law(parabola(at_eutectoid), g, undercooling);
        out[0] = f.A;
        out[1] = f.B;
        out[2] = f.C;
        out[3] = f.D;
        return 0;
    } catch (const std::exception& e) {
        return fail(err, err_len, e.what());
    }
}

extern "C" int pearl_phase_diagram_law(const double at_eutectoid[12], const double c_eq[3], const double slope[4],
                                       double undercooling, double out[12], char* err, int err_len) {
    if (!at_eutectoid || !c_eq || !slope || !out) return fail(err, err_len, "invalid argument");
    try {
        pearl_host::Thermo t;
        for (int p = 0; p < 3; ++p) t[p] = parabola(at_eutectoid + 4 * p);
        pearl_host::PhaseDiagram pd;
        for (int k = 0; k < 3; ++k) pd.c_eq[k] = c_eq[k];
        for (int k = 0; k < 4; ++k) pd.slope[k] = slope[k];
        const pearl_host::Thermo r = pearl_host::phase_diagram_law(t, pd, undercooling);
        for (int p = 0; p < 3; ++p) {
            out[4 * p] = r[p].A;
            out[4 * p + 1] = r[p].B;
            out[4 * p + 2] = r[p].C;
            out[4 * p + 3] = r[p].D;
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(err, err_len, e.what());
    }
}
