// pearl_gpu: the host program around the MI355X step.  A drop-in for the
// reference's pearl_serial on the same inputs: same command line
// (pearl_serial.cpp:2190-2206), strict input parser (:1600-1758), checkpoint
// signature (:1566-1597), initialization (:1761-1837), HDF5/XMF frames and
// checkpoints (:1914-2020), diagnostics.csv (:2263, :2288) and status files
// (:2319-2322).  The step loop (:2296-2316) calls the C ABI of
// libpearl_gpu.so (include/pearl_gpu.h) for advance, the moving window and
// the snapshot reductions; every field update of the loop runs on the GPU.
//
// Initialization (filling, the capillarity-only pre-smoothing and the carbon
// set-up) is host preprocessing, as in the reference, and runs once on the
// CPU before the state is uploaded.
#include <hdf5.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <csignal>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <limits>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "pearl_gpu.h"

namespace host {

// The reference's checkpoint signature names the inherited numerical model by
// this digest (pearl_serial.cpp:31-33); checkpoints are only interchangeable
// with the reference when the signatures are equal, so the same value is used.
constexpr const char* kModelDigest = "344eef3b2e1fdbd8a41d81378491b923f6d815d8ef917f8397729c5a57965580";

using Three = std::array<double, 3>;

struct Parabola {
    double A = 0, B = 0, C = 0, D = 0;
    void validate() const {
        if (!std::isfinite(A) || !std::isfinite(B) || !std::isfinite(C) || !std::isfinite(D) || A <= 0.0 || D < 0.0)
            throw std::invalid_argument("Parabola requires finite A>0, B, C, D>=0");
    }
    double mu(double c) const { return std::fma(2.0 * A, c, B); }
    double c_at_mu(double m) const { return (m - B) / (2.0 * A); }
    double chi() const { return 1.0 / (2.0 * A); }
};
using Thermo = std::array<Parabola, 3>;

// ---------------------------------------------------------------------------
// Temperature laws (pearl_serial.cpp:186-276), long double as the reference.
// ---------------------------------------------------------------------------
struct GP {
    long double q2 = 0, q1 = 0, q0 = 0;
};

GP to_gp(const Parabola& p) {
    p.validate();
    const long double A = p.A, B = p.B, C = p.C;
    return {-1 / (4 * A), B / (2 * A), C - B * B / (4 * A)};
}

Parabola from_gp(const GP& g, double D) {
    if (!std::isfinite(g.q2) || !std::isfinite(g.q1) || !std::isfinite(g.q0) || g.q2 >= 0)
        throw std::invalid_argument("temperature law requires finite q2<0, q1, q0");
    const long double A = -1 / (4 * g.q2), B = -g.q1 / (2 * g.q2);
    Parabola p{static_cast<double>(A), static_cast<double>(B), static_cast<double>(g.q0 + B * B / (4 * A)), D};
    p.validate();
    return p;
}

Parabola linear_gp(const Parabola& ref, const GP& slope, double undercooling) {
    if (!std::isfinite(undercooling) || undercooling < 0)
        throw std::invalid_argument("undercooling must be finite and nonnegative");
    if (!std::isfinite(slope.q2) || !std::isfinite(slope.q1) || !std::isfinite(slope.q0))
        throw std::invalid_argument("nonfinite temperature derivative");
    ref.validate();
    if (undercooling == 0) return ref;
    GP g = to_gp(ref);
    const long double dT = -static_cast<long double>(undercooling);
    g.q2 += dT * slope.q2;
    g.q1 += dT * slope.q1;
    g.q0 += dT * slope.q0;
    return from_gp(g, ref.D);
}

struct PhaseDiagram {
    Three ceq{};
    std::array<double, 4> slope{};
};

Thermo phase_diagram_law(const Thermo& ref, const PhaseDiagram& pd, double undercooling) {
    if (!std::isfinite(undercooling) || undercooling < 0)
        throw std::invalid_argument("undercooling must be finite and nonnegative");
    for (const auto& p : ref) p.validate();
    for (double c : pd.ceq)
        if (!std::isfinite(c) || c < 0 || c > 1) throw std::invalid_argument("phase-diagram reference carbon must be in [0,1]");
    for (double s : pd.slope)
        if (!std::isfinite(s)) throw std::invalid_argument("phase-diagram slope must be finite dc/dT");
    const long double dT = -static_cast<long double>(undercooling);
    const long double gA = ref[2].A, gc = pd.ceq[2];
    Thermo out = ref;
    for (int p = 0; p <= 1; ++p) {
        const long double pA = ref[p].A, pc = pd.ceq[p];
        const long double dp = dT * pd.slope[2 * p], dg = dT * pd.slope[2 * p + 1];
        const long double cp = pc + dp, cg = gc + dg;
        if (!std::isfinite(cp) || !std::isfinite(cg) || cp < 0 || cp > 1 || cg < 0 || cg > 1)
            throw std::invalid_argument("phase-diagram extrapolated carbon outside [0,1]");
        if (undercooling == 0) continue;
        out[p].B = static_cast<double>(ref[p].B + 2 * (gA * dg - pA * dp));
        out[p].C = static_cast<double>(ref[p].C + pA * dp * (2 * pc + dp) - gA * dg * (2 * gc + dg));
        out[p].validate();
    }
    return out;
}

// ---------------------------------------------------------------------------
// Filling recipes (pearl_serial.cpp:1201-1514)
// ---------------------------------------------------------------------------
struct Fill {
    std::string method;
    std::vector<double> p;
};

std::string trim(const std::string& s) {
    const auto a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return {};
    return s.substr(a, s.find_last_not_of(" \t\r\n") - a + 1);
}

Fill parse_fill(const std::string& method, const std::string& value, int line) {
    auto bad = [&]() {
        return std::invalid_argument("Invalid " + method + " filling parameters at input line " + std::to_string(line) +
                                     ": expected {finite_number,...}, optionally followed by ';'");
    };
    std::string lst = trim(value);
    if (!lst.empty() && lst.back() == ';') lst = trim(lst.substr(0, lst.size() - 1));
    if (lst.size() < 3 || lst.front() != '{' || lst.back() != '}') throw bad();
    lst = lst.substr(1, lst.size() - 2);
    Fill f{method, {}};
    std::size_t start = 0;
    while (true) {
        const std::size_t comma = lst.find(',', start);
        const std::string item = trim(lst.substr(start, comma == std::string::npos ? std::string::npos : comma - start));
        if (item.empty()) throw bad();
        std::size_t used = 0;
        double v = 0;
        try {
            v = std::stod(item, &used);
        } catch (const std::exception&) {
            throw bad();
        }
        if (used != item.size() || !std::isfinite(v)) throw bad();
        f.p.push_back(v);
        if (comma == std::string::npos) break;
        start = comma + 1;
    }
    return f;
}

void validate_fills(const std::vector<Fill>& fills, int nx, int ny) {
    if (nx <= 0 || ny <= 0) throw std::invalid_argument("Filling requires positive nx and ny");
    for (const auto& f : fills) {
        const auto& p = f.p;
        auto fail = [&](const std::string& why) { throw std::invalid_argument(f.method + ": " + why); };
        auto count = [&](std::size_t n) {
            if (p.size() != n) fail("expected " + std::to_string(n) + " parameters");
            for (double v : p)
                if (!std::isfinite(v)) fail("parameters must be finite");
        };
        auto integer = [&](std::size_t i) {
            if (std::trunc(p[i]) != p[i] || p[i] < static_cast<double>(std::numeric_limits<int>::min()) ||
                p[i] > static_cast<double>(std::numeric_limits<int>::max()))
                fail("parameter " + std::to_string(i + 1) + " must be a representable integer");
        };
        auto phase = [&](std::size_t i) {
            integer(i);
            if (p[i] < 0 || p[i] > 2)
                fail("phase IDs must be alpha=0, theta=1 or gamma=2; gamma grain clones are unsupported");
        };
        auto distinct = [&](std::initializer_list<std::size_t> ids) {
            for (auto i : ids) phase(i);
            for (auto a = ids.begin(); a != ids.end(); ++a)
                for (auto b = a + 1; b != ids.end(); ++b)
                    if (p[*a] == p[*b]) fail("listed phase IDs must be distinct");
        };
        auto half_open = [&]() {
            for (std::size_t i = 0; i < 4; ++i) integer(i);
            if (!(0 <= p[0] && p[0] < p[1] && p[1] <= nx && 0 <= p[2] && p[2] < p[3] && p[3] <= ny))
                fail("box must satisfy 0<=xlo<xhi<=nx and 0<=ylo<yhi<=ny");
        };
        auto center = [&](std::size_t x, std::size_t y, bool ints) {
            if (ints) {
                integer(x);
                integer(y);
            }
            if (!(0 <= p[x] && p[x] < nx && 0 <= p[y] && p[y] < ny)) fail("center/junction must be inside the physical grid");
        };
        auto positive = [&](std::size_t i) {
            if (!(p[i] > 0.0)) fail("parameter " + std::to_string(i + 1) + " must be positive");
        };
        auto fraction = [&](std::size_t i) {
            if (!(p[i] > 0.0 && p[i] < 1.0)) fail("cementite fraction must be strictly between 0 and 1");
        };
        auto nonneg = [&](std::size_t i) {
            if (p[i] < 0.0) fail("parameter " + std::to_string(i + 1) + " must be nonnegative");
        };
        auto angle = [&](std::size_t i) {
            if (!std::isfinite(p[i] * 3.14159265358979323846 / 180.0)) fail("angle conversion to radians overflows");
        };
        auto sq = [&](std::size_t i) {
            const double s = p[i] * p[i];
            if (!std::isfinite(s) || (p[i] > 0.0 && s == 0.0)) fail("length squared overflows or underflows");
        };
        auto resolved = [&](double v, const std::string& name) {
            if (!(v > 0.0) || !std::isfinite(v)) fail(name + " must remain finite and positive");
        };
        const std::string& m = f.method;
        if (m == "FILLCUBE") {
            count(7);
            phase(0);
            for (std::size_t i = 1; i < 7; ++i) integer(i);
            if (p[3] != 0 || p[6] != 0) fail("2D filling requires zlo=zhi=0");
            if (!(0 <= p[1] && p[1] <= p[4] && p[4] < nx && 0 <= p[2] && p[2] <= p[5] && p[5] < ny))
                fail("inclusive bounds must satisfy 0<=xlo<=xhi<nx and 0<=ylo<=yhi<ny");
        } else if (m == "FILLCENTERBOX2D") {
            count(6);
            half_open();
            distinct({4, 5});
        } else if (m == "FILLYJUNCTION2D" || m == "FILLYJUNCTIONLAMELLAE2D") {
            const bool lam = m == "FILLYJUNCTIONLAMELLAE2D";
            count(lam ? 16 : 13);
            half_open();
            center(4, 5, true);
            positive(9);
            positive(10);
            angle(6);
            angle(7);
            angle(8);
            if (lam) {
                positive(11);
                positive(12);
                distinct({13, 14, 15});
                resolved(p[11] + p[12], "lamellar period");
            } else {
                distinct({11, 12});
            }
        } else if (m == "FILLCYLINDER") {
            count(6);
            phase(0);
            center(1, 2, true);
            integer(3);
            integer(4);
            positive(5);
            if (p[3] != 0 || p[4] != 0) fail("2D cylinder cross-section requires zlo=zhi=0");
            sq(5);
        } else if (m == "FILLADHEREDPEARLITEARM2D") {
            count(11);
            center(0, 1, false);
            positive(3);
            integer(4);
            positive(4);
            if (p[4] > std::numeric_limits<int>::max() / 2) fail("pair_count is too large");
            fraction(5);
            positive(6);
            nonneg(7);
            nonneg(8);
            distinct({9, 10});
            if (p[7] > p[6]) fail("contact_half_width must not exceed peak_half_width");
            angle(2);
            sq(3);
            sq(6);
            sq(8);
            const double period = p[3] / p[4];
            resolved(period, "arm period");
            resolved((1.0 - p[5]) * period, "ferrite segment length");
            resolved(p[5] * period, "cementite segment length");
        } else if (m == "FILLPEARLITEELLIPSE2D") {
            count(11);
            center(0, 1, false);
            nonneg(3);
            integer(4);
            positive(4);
            positive(5);
            positive(6);
            fraction(7);
            nonneg(8);
            distinct({9, 10});
            angle(2);
            sq(8);
            const double length = p[4] * p[5];
            const double semi_major = 0.5 * length, semi_minor = p[6] * semi_major;
            resolved(length, "ellipse length");
            resolved(semi_major, "ellipse semimajor axis");
            resolved(semi_minor, "ellipse semiminor axis");
            resolved(p[7] * p[5], "cementite lamellar width");
            if (!std::isfinite((p[3] + length) * (p[3] + length))) fail("ellipse endpoint distance overflows");
        } else if (m == "FILLTHREEGRAINYJUNCTION2D" || m == "FILLYJUNCTIONPEARLITE2D") {
            fail("historical multi-grain initialization requires separate gamma grain fields; "
                 "Pearl has one gamma field and will not merge grain IDs silently");
        } else {
            fail("unsupported filling method in the 2D three-phase Pearl solver; see FILLING_METHODS.md");
        }
    }
}

Three one_hot(int phase) {
    Three t{0.0, 0.0, 0.0};
    t[phase] = 1.0;
    return t;
}

int centered_lamella(double along, double period, double frac, int ferrite, int cementite) {
    double pos = std::fmod(along, period);
    if (pos < 0.0) pos += period;
    const double w = frac * period;
    const double a = 0.5 * (period - w), b = 0.5 * (period + w);
    return pos >= a && pos < b ? cementite : ferrite;
}

Three fill_phase_at(const std::vector<Fill>& fills, int gx, int gy) {
    Three out = one_hot(2);
    const double pi = 3.14159265358979323846;
    for (const auto& f : fills) {
        const auto& p = f.p;
        const std::string& m = f.method;
        if (m == "FILLCUBE") {
            const int sel = static_cast<int>(p[0]);
            if (sel != 2 && gx >= p[1] && gx <= p[4] && gy >= p[2] && gy <= p[5]) out = one_hot(sel);
        } else if (m == "FILLCYLINDER") {
            const double dx = gx - p[1], dy = gy - p[2];
            const int sel = static_cast<int>(p[0]);
            if (sel != 2 && dx * dx + dy * dy <= p[5] * p[5]) out = one_hot(sel);
        } else if (m == "FILLCENTERBOX2D") {
            const bool in = gx >= p[0] && gx < p[1] && gy >= p[2] && gy < p[3];
            out = one_hot(static_cast<int>(in ? p[4] : p[5]));
        } else if (m == "FILLYJUNCTION2D" || m == "FILLYJUNCTIONLAMELLAE2D") {
            const bool lam = m == "FILLYJUNCTIONLAMELLAE2D";
            const int ferrite = static_cast<int>(p[lam ? 13 : 11]), aust = static_cast<int>(p[lam ? 15 : 12]);
            out = one_hot(aust);
            if (!(gx >= p[0] && gx < p[1] && gy >= p[2] && gy < p[3])) continue;
            const double dx = gx - p[4], dy = gy - p[5];
            for (int arm = 0; arm < 3; ++arm) {
                const double ang = p[6 + arm] * pi / 180.0;
                const double ux = std::cos(ang), uy = std::sin(ang);
                const double along = dx * ux + dy * uy, across = -dx * uy + dy * ux;
                if (along >= 0.0 && along <= p[9] && std::fabs(across) <= p[10]) {
                    int sel = ferrite;
                    if (lam) {
                        const double period = p[11] + p[12];
                        double pos = std::fmod(along, period);
                        if (pos < 0.0) pos += period;
                        if (!(pos < p[11])) sel = static_cast<int>(p[14]);
                    }
                    out = one_hot(sel);
                    break;
                }
            }
        } else if (m == "FILLADHEREDPEARLITEARM2D") {
            const double ang = p[2] * pi / 180.0;
            const double ux = std::cos(ang), uy = std::sin(ang);
            const double dx = gx - p[0], dy = gy - p[1];
            const double along = dx * ux + dy * uy, across = -dx * uy + dy * ux;
            const double len = p[3];
            const int pairs = static_cast<int>(p[4]);
            const double period = len / static_cast<double>(pairs);
            const double fw = (1.0 - p[5]) * period, cw = p[5] * period;
            const double peak = p[6], contact = p[7], bridge = p[8];
            const int ferrite = static_cast<int>(p[9]), cementite = static_cast<int>(p[10]);
            int sel = -1;
            if (along >= 0.0 && along <= len) {
                int pid = along >= len ? pairs - 1 : static_cast<int>(std::floor(along / period));
                pid = std::max(0, std::min(pairs - 1, pid));
                const double start = static_cast<double>(pid) * period;
                const double local = along - start;
                int seg;
                double seg_start, seg_len;
                if (local < fw) {
                    seg = 2 * pid;
                    seg_start = start;
                    seg_len = fw;
                    sel = ferrite;
                } else {
                    seg = 2 * pid + 1;
                    seg_start = start + fw;
                    seg_len = cw;
                    sel = cementite;
                }
                double t = (along - seg_start) / seg_len;
                if (t < 0.0) t = 0.0;
                else if (t > 1.0) t = 1.0;
                const double h0 = seg == 0 ? 0.0 : contact;
                const double h1 = seg == 2 * pairs - 1 ? 0.0 : contact;
                const double avg = 0.5 * (h0 + h1);
                double half = (1.0 - t) * h0 + t * h1 + (peak - avg) * std::sin(pi * t);
                if (half > peak) half = peak;
                if (std::fabs(across) > half) sel = -1;
            }
            if (bridge > 0.0) {
                const double b2 = bridge * bridge;
                const double d0 = along * along + across * across;
                const double d1 = (along - len) * (along - len) + across * across;
                if (d0 <= b2) sel = ferrite;
                if (d1 <= b2) sel = cementite;
            }
            if (sel >= 0) out = one_hot(sel);
        } else if (m == "FILLPEARLITEELLIPSE2D") {
            const double ang = p[2] * pi / 180.0;
            const double ux = std::cos(ang), uy = std::sin(ang);
            const double dx = gx - p[0], dy = gy - p[1];
            const double along = dx * ux + dy * uy, across = -dx * uy + dy * ux;
            const double start = p[3], period = p[5], length = p[4] * period;
            const double semi_major = 0.5 * length, semi_minor = p[6] * semi_major;
            const double center = start + semi_major;
            const double na = (along - center) / semi_major, nc = across / semi_minor;
            const double bridge = p[8];
            const int ferrite = static_cast<int>(p[9]), cementite = static_cast<int>(p[10]);
            int sel = -1;
            if (na * na + nc * nc <= 1.0) sel = centered_lamella(along - start, period, p[7], ferrite, cementite);
            if (bridge > 0.0) {
                const double d0 = (along - start) * (along - start) + across * across;
                const double d1 = (along - (start + length)) * (along - (start + length)) + across * across;
                if (d0 <= bridge * bridge || d1 <= bridge * bridge) sel = ferrite;
            }
            if (sel >= 0) out = one_hot(sel);
        } else {
            throw std::invalid_argument("Unvalidated/unsupported filling method: " + m);
        }
    }
    return out;
}

// ---------------------------------------------------------------------------
// Config (pearl_serial.cpp:1544-1758)
// ---------------------------------------------------------------------------
struct Config {
    int nx = 0, ny = 0, theta_columns = 0, phase_a = 0, phase_b = 2;
    double dx = 0, dy = 0, end_time = 0, output_dt = 0, seed_height = 0, radius = 0;
    double initial_mu = 0, initial_carbon = 0;
    std::string composition_mode, initialization = "lamella", mode = "coupled";
    int smooth_steps = 100, smooth_min_steps = 100;
    double smooth_energy_rtol = 1e-10, smooth_max_phase_change = 0.02;
    Thermo thermo{}, reference_thermo{};
    std::array<GP, 3> slope{};
    std::string thermo_mode = "direct_at_temperature", temperature_law = "linear_gp";
    PhaseDiagram diagram;
    double T = 0, T_eutectoid = 0, undercooling = 0, max_undercooling = 0;
    double eps = 0, sigma = 0, triple = 0, Vm = 0;
    Three tau{};
    pearl_params ctl{};  // numerical controls in the C ABI layout
    double dt = 0, shift_trigger = 0, shift_target = 0, shift_gamma_carbon = 0;
    long long max_steps = 100000000;
    bool moving_window = false;
    std::vector<Fill> fills;
    std::map<std::string, std::string> raw;

    std::string physics_signature() const {
        std::ostringstream o;
        o << std::setprecision(17) << "pearlite-v0.4-reference;source=" << kModelDigest
          << ";3ph;binary;equal-sigma;Jat=0;Xperiodic;Ynoflux;";
        o << nx << ';' << ny << ';' << dx << ';' << dy << ';' << Vm << ';' << eps << ';' << sigma << ';' << triple << ';';
        o << thermo_mode << ';' << T << ';' << T_eutectoid << ';' << undercooling << ';' << max_undercooling << ';';
        if (thermo_mode == "eutectoid_reference") {
            for (const auto& p : reference_thermo) o << p.A << ';' << p.B << ';' << p.C << ';';
            if (temperature_law == "linear_gp") {
                for (const auto& s : slope) o << s.q2 << ';' << s.q1 << ';' << s.q0 << ';';
            } else {
                o << "phase-diagram-fixed-curvature-v1;";
                for (double c : diagram.ceq) o << c << ';';
                for (double s : diagram.slope) o << s << ';';
            }
        }
        for (const auto& p : thermo) o << p.A << ';' << p.B << ';' << p.C << ';' << p.D << ';';
        for (double t : tau) o << t << ';';
        o << mode << ';' << initialization << ';' << seed_height << ';' << theta_columns << ';' << radius << ';'
          << phase_a << ';' << phase_b << ';' << composition_mode << ';' << initial_mu << ';' << initial_carbon << ';';
        o << smooth_steps << ';' << smooth_min_steps << ';' << smooth_energy_rtol << ';' << smooth_max_phase_change << ';';
        o << moving_window << ';' << shift_trigger << ';' << shift_target << ';' << shift_gamma_carbon;
        if (initialization == "microsim") {
            o << ";filling-v1";
            for (const auto& f : fills) {
                o << ';' << f.method;
                for (double v : f.p) o << ';' << v;
            }
        }
        return o.str();
    }
};

Config read_config(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open input: " + path);
    Config c;
    std::string line;
    int ln = 0;
    while (std::getline(in, line)) {
        ++ln;
        line = trim(line.substr(0, line.find('#')));
        if (line.empty()) continue;
        const auto pos = line.find('=');
        if (pos == std::string::npos) throw std::runtime_error("missing = on line " + std::to_string(ln));
        const std::string key = trim(line.substr(0, pos)), val = trim(line.substr(pos + 1));
        if (key.rfind("FILL", 0) == 0) {
            c.fills.push_back(parse_fill(key, val, ln));
            continue;
        }
        if (key.empty() || val.empty() || c.raw.count(key)) throw std::runtime_error("empty/duplicate key: " + key);
        c.raw[key] = val;
    }
    std::set<std::string> used;
    auto str = [&](const std::string& k, const std::string& def = "") {
        auto it = c.raw.find(k);
        if (it == c.raw.end()) return def;
        used.insert(k);
        return it->second;
    };
    auto number = [&](const std::string& k, double def = std::numeric_limits<double>::quiet_NaN()) {
        const std::string s = str(k);
        if (s.empty()) {
            if (!std::isfinite(def)) throw std::runtime_error("missing required key: " + k);
            return def;
        }
        std::size_t n = 0;
        const double v = std::stod(s, &n);
        if (n != s.size() || !std::isfinite(v)) throw std::runtime_error("invalid number: " + k);
        return v;
    };
    auto integer = [&](const std::string& k, int def = -1) {
        const double a = number(k, def < 0 ? std::numeric_limits<double>::quiet_NaN() : def);
        if (a != std::floor(a) || std::abs(a) > 1e8) throw std::runtime_error("invalid integer: " + k);
        return static_cast<int>(a);
    };
    c.nx = integer("nx");
    c.ny = integer("ny");
    c.dx = number("dx");
    c.dy = number("dy", c.dx);
    c.eps = number("epsilon");
    c.sigma = number("sigma");
    c.triple = number("triple");
    c.Vm = number("Vm");
    c.tau = {number("tau_at"), number("tau_ag"), number("tau_tg")};
    const std::array<std::string, 3> names{"alpha", "theta", "gamma"};
    c.thermo_mode = str("thermo_mode");
    if (c.thermo_mode == "eutectoid_reference") {
        c.T_eutectoid = number("T_eutectoid");
        c.undercooling = number("undercooling");
        c.max_undercooling = number("max_undercooling");
        c.temperature_law = str("temperature_law");
        if (c.temperature_law != "linear_gp" && c.temperature_law != "phase_diagram_fixed_curvature")
            throw std::runtime_error("temperature_law must be linear_gp or phase_diagram_fixed_curvature");
        c.T = c.T_eutectoid - c.undercooling;
        if (c.T_eutectoid <= 0 || c.T <= 0 || c.undercooling < 0 || c.max_undercooling < 0 ||
            c.undercooling > c.max_undercooling)
            throw std::runtime_error("temperature outside declared undercooling interval");
        for (int i = 0; i < 3; ++i) {
            const auto& n = names[i];
            c.reference_thermo[i] = {number("Aeq_" + n), number("Beq_" + n), number("Ceq_" + n), number("D_" + n)};
            if (c.temperature_law == "linear_gp") {
                c.slope[i] = {number("dq2_dT_" + n), number("dq1_dT_" + n), number("dq0_dT_" + n)};
                c.thermo[i] = linear_gp(c.reference_thermo[i], c.slope[i], c.undercooling);
            }
        }
        if (c.temperature_law == "phase_diagram_fixed_curvature") {
            c.diagram.ceq = {number("ceq_alpha"), number("ceq_theta"), number("ceq_gamma")};
            c.diagram.slope = {number("slope_alpha_on_alpha_gamma"), number("slope_gamma_on_alpha_gamma"),
                               number("slope_theta_on_theta_gamma"), number("slope_gamma_on_theta_gamma")};
            c.thermo = phase_diagram_law(c.reference_thermo, c.diagram, c.undercooling);
        }
    } else if (c.thermo_mode == "direct_at_temperature") {
        c.T = number("temperature");
        if (c.T <= 0) throw std::runtime_error("temperature must be positive");
        for (int i = 0; i < 3; ++i) {
            const auto& n = names[i];
            c.thermo[i] = {number("A_" + n), number("B_" + n), number("C_" + n), number("D_" + n)};
            c.thermo[i].validate();
        }
    } else {
        throw std::runtime_error(
            "Set thermo_mode=eutectoid_reference or direct_at_temperature. "
            "Legacy inputs require an explicit mode and temperature; no temperature is guessed.");
    }
    pearl_params& k = c.ctl;
    k.dt_max = 1e-3;
    k.dt_min = 1e-12;
    k.max_phase_change = .04;
    k.cg_rtol = 1e-11;
    k.cg_atol = 1e-14;
    k.closure_tol = 1e-10;
    k.energy_rtol = 1e-11;
    k.energy_atol = 1e-12;
    k.cg_max = 2000;
    k.max_retries = 24;
    k.check_energy = 1;
    c.dt = number("dt");
    k.dt_max = number("dt_max", c.dt);
    k.dt_min = number("dt_min", c.dt / 1048576.0);
    c.end_time = number("end_time");
    c.output_dt = number("output_dt");
    c.max_steps = integer("max_steps", 100000000);
    c.initialization = str("initialization", "lamella");
    c.mode = str("mode", "coupled");
    if (c.mode != "coupled" && c.mode != "phase_only") throw std::runtime_error("mode must be coupled or phase_only");
    k.phase_only = c.mode == "phase_only";
    c.seed_height = number("seed_height", c.ny * c.dy / 4);
    c.theta_columns = integer("theta_columns", c.nx / 2);
    c.radius = number("radius", std::min(c.nx * c.dx, c.ny * c.dy) / 4);
    c.phase_a = integer("phase_a", 0);
    c.phase_b = integer("phase_b", 2);
    if (c.initialization != "lamella" && c.initialization != "sharp_smooth" && c.initialization != "flat" &&
        c.initialization != "circle" && c.initialization != "microsim")
        throw std::runtime_error("unknown initialization");
    if (c.initialization == "microsim") {
        if (c.fills.empty()) throw std::runtime_error("initialization=microsim requires at least one FILL operation");
        validate_fills(c.fills, c.nx, c.ny);
    } else if (!c.fills.empty()) {
        throw std::runtime_error("FILL operations require initialization=microsim; no geometry is silently ignored");
    }
    c.smooth_steps = integer("smooth_steps", 100);
    c.smooth_min_steps = integer("smooth_min_steps", c.smooth_steps);
    c.smooth_energy_rtol = number("smooth_energy_rtol", 1e-10);
    c.smooth_max_phase_change = number("smooth_max_phase_change", 0.02);
    if (c.phase_a < 0 || c.phase_a > 2 || c.phase_b < 0 || c.phase_b > 2 || c.phase_a == c.phase_b)
        throw std::runtime_error("invalid phase pair");
    const int choices = static_cast<int>(c.raw.count("initial_mu") + c.raw.count("gamma_carbon") + c.raw.count("mean_carbon"));
    if (choices != 1) throw std::runtime_error("specify exactly one of initial_mu, gamma_carbon, mean_carbon");
    if (c.raw.count("initial_mu")) {
        c.composition_mode = "mu";
        c.initial_mu = number("initial_mu");
    }
    if (c.raw.count("gamma_carbon")) {
        c.composition_mode = "gamma";
        c.initial_carbon = number("gamma_carbon");
        c.initial_mu = c.thermo[2].mu(c.initial_carbon);
    }
    if (c.raw.count("mean_carbon")) {
        c.composition_mode = "mean";
        c.initial_carbon = number("mean_carbon");
    }
    k.max_phase_change = number("max_phase_change", .04);
    k.cg_rtol = number("cg_rtol", 1e-11);
    k.cg_atol = number("cg_atol", 1e-14);
    k.cg_max = integer("cg_max", 2000);
    k.closure_tol = number("closure_tol", 1e-10);
    k.energy_atol = number("energy_atol", 1e-13 * c.sigma * std::max(c.nx * c.dx, c.ny * c.dy));
    c.moving_window = integer("moving_window", 0) != 0;
    c.shift_trigger = number("shift_trigger", .65 * c.ny * c.dy);
    c.shift_target = number("shift_target", .35 * c.ny * c.dy);
    c.shift_gamma_carbon =
        number("shift_gamma_carbon", c.composition_mode == "gamma" ? c.initial_carbon : c.thermo[2].c_at_mu(c.initial_mu));
    for (const auto& kv : c.raw)
        if (!used.count(kv.first)) throw std::runtime_error("unknown key: " + kv.first);
    if (!(c.eps > 0) || !(c.sigma > 0) || c.triple < 0 || !std::isfinite(c.eps) || !std::isfinite(c.sigma) ||
        !std::isfinite(c.triple))
        throw std::invalid_argument("Invalid interface parameters");
    for (double t : c.tau)
        if (!(t > 0) || !std::isfinite(t)) throw std::invalid_argument("tau must be finite and positive");
    if (!(c.Vm > 0) || !std::isfinite(c.Vm)) throw std::invalid_argument("Invalid Vm");
    if (c.nx < 3 || c.ny < 2 || c.dx <= 0 || c.dy <= 0 || c.dt <= 0 || k.dt_min <= 0 || k.dt_max < c.dt ||
        k.dt_min > c.dt || c.end_time <= 0 || c.output_dt <= 0 || c.max_steps < 1)
        throw std::runtime_error("invalid geometry/time bounds");
    if (c.theta_columns < 1 || c.theta_columns >= c.nx || c.seed_height <= 0 || c.seed_height >= c.ny * c.dy)
        throw std::runtime_error("invalid seed geometry");
    if (k.max_phase_change <= 0 || k.max_phase_change > .2 || k.cg_rtol <= 0 || k.cg_atol <= 0 || k.closure_tol <= 0 ||
        k.energy_atol < 0)
        throw std::runtime_error("invalid numerical control");
    if (c.smooth_steps < 1 || c.smooth_min_steps < 0 || c.smooth_min_steps > c.smooth_steps ||
        c.smooth_energy_rtol <= 0 || c.smooth_max_phase_change <= 0 || c.smooth_max_phase_change > .2)
        throw std::runtime_error("invalid smoothing control");
    if (c.moving_window && (c.shift_target <= 0 || c.shift_trigger <= c.shift_target || c.shift_trigger >= c.ny * c.dy ||
                            c.shift_gamma_carbon < 0 || c.shift_gamma_carbon > 1))
        throw std::runtime_error("invalid moving window settings");
    // C ABI parameter block
    k.nx = c.nx;
    k.ny = c.ny;
    k.dx = c.dx;
    k.dy = c.dy;
    for (int i = 0; i < 3; ++i) {
        k.A[i] = c.thermo[i].A;
        k.B[i] = c.thermo[i].B;
        k.C[i] = c.thermo[i].C;
        k.D[i] = c.thermo[i].D;
        k.tau[i] = c.tau[i];
    }
    k.eps = c.eps;
    k.sigma = c.sigma;
    k.triple = c.triple;
    k.Vm = c.Vm;
    return c;
}

// ---------------------------------------------------------------------------
// Host fields: planar [3][ny][nx] phi, [ny][nx] carbon and mu.
// ---------------------------------------------------------------------------
struct Fields {
    int nx = 0, ny = 0;
    std::vector<double> phi, c, mu;
    std::size_t n() const { return static_cast<std::size_t>(nx) * ny; }
    Three cell(std::size_t i) const { return {phi[i], phi[n() + i], phi[2 * n() + i]}; }
    void set(std::size_t i, const Three& p) {
        for (int a = 0; a < 3; ++a) phi[a * n() + i] = p[a];
    }
};

struct Mix {
    double s = 0, c0 = 0, K = 0;
};

Mix mixture(const Three& phi, const Thermo& f) {  // :125-138
    const double trip = 2.0 * phi[0] * phi[1] * phi[2];
    Mix m;
    for (int p = 0; p < 3; ++p) {
        const double h = phi[p] * phi[p] * (3.0 - 2.0 * phi[p]) + trip;
        const double chi = f[p].chi();
        m.s += h * chi;
        m.c0 -= h * chi * f[p].B;
        m.K += phi[p] * f[p].D * chi;
    }
    if (!(m.s > 0.0) || !std::isfinite(m.s) || !std::isfinite(m.c0) || !std::isfinite(m.K) || m.K < 0.0)
        throw std::runtime_error("Invalid mixture coefficients");
    return m;
}

double profile(double distance, double eps) {  // :563-569
    const double pi = std::acos(-1.0), half = pi * pi * eps / 8;
    if (distance <= -half) return 1;
    if (distance >= half) return 0;
    return 0.5 * (1 - std::sin(4 * distance / (pi * eps)));
}

Three project_simplex(const Three& z) {  // :491-509
    for (double v : z)
        if (!std::isfinite(v)) throw std::invalid_argument("Nonfinite projection input");
    Three s = z;
    std::sort(s.begin(), s.end(), std::greater<double>());
    double sum = 0.0, thr = 0.0;
    int active = 0;
    for (int j = 0; j < 3; ++j) {
        sum += s[j];
        const double cand = (sum - 1.0) / (j + 1);
        if (s[j] > cand) {
            active = j + 1;
            thr = cand;
        }
    }
    if (active == 0) throw std::runtime_error("Simplex projection failed");
    Three out{};
    for (int a = 0; a < 3; ++a) out[a] = std::max(0.0, z[a] - thr);
    return out;
}

// interface_energy_only :822-841 (sequential order)
double interface_energy(const Fields& F, const Config& c) {
    const double obs = 16 * c.sigma / (std::acos(-1.) * std::acos(-1.) * c.eps), tr = c.triple / c.eps;
    const int nx = F.nx, ny = F.ny;
    double total = 0;
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const std::size_t i = static_cast<std::size_t>(y) * nx + x;
            const Three p = F.cell(i);
            total += obs * (p[0] * p[1] + p[0] * p[2] + p[1] * p[2]) + tr * p[0] * p[1] * p[2];
            for (int dir = 0; dir < 2; ++dir) {
                if (dir == 1 && y == ny - 1) continue;
                const std::size_t j = dir == 0 ? static_cast<std::size_t>(y) * nx + (x == nx - 1 ? 0 : x + 1) : i + nx;
                const Three q = F.cell(j);
                const double dist = dir == 0 ? c.dx : c.dy;
                for (int a = 0; a < 3; ++a)
                    for (int b = a + 1; b < 3; ++b) {
                        const double z = (p[a] * q[b] - q[a] * p[b]) / dist;
                        total += c.eps * c.sigma * z * z;
                    }
            }
        }
    return total * c.dx * c.dy;
}

struct SmoothReport {
    int steps = 0;
    bool converged = false;
    double e0 = 0, e1 = 0, rel = 0, change = 0;
    Three before{}, after{};
};

Three phase_fractions(const Fields& F) {
    Three s{0, 0, 0};
    for (std::size_t i = 0; i < F.n(); ++i)
        for (int a = 0; a < 3; ++a) s[a] += F.phi[a * F.n() + i];
    const double n = static_cast<double>(F.nx) * F.ny;
    for (auto& v : s) v = v / n;
    return s;
}

// smooth_interface_only :869-958
SmoothReport smooth(Fields& F, const Config& c) {
    if (c.smooth_steps < 1 || c.smooth_min_steps < 0 || c.smooth_min_steps > c.smooth_steps || !(c.smooth_energy_rtol > 0) ||
        !(c.smooth_max_phase_change > 0) || c.smooth_max_phase_change > .2)
        throw std::invalid_argument("invalid sharp-smoothing controls");
    SmoothReport rep;
    rep.before = phase_fractions(F);
    rep.e0 = interface_energy(F, c);
    double previous = rep.e0;
    const double eps = c.eps, sigma = c.sigma;
    const double obs = 16 * sigma / (std::acos(-1.) * std::acos(-1.) * eps), tr = c.triple / eps;
    const double base = .12 * std::min(c.dx * c.dx, c.dy * c.dy) / (eps * sigma);
    const int nx = F.nx, ny = F.ny;
    const std::size_t n = F.n();
    std::vector<double> force(3 * n);
    Fields trial = F;
    for (int step = 1; step <= c.smooth_steps; ++step) {
        double fmax = 0;
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const std::size_t i = static_cast<std::size_t>(y) * nx + x;
                const Three p = F.cell(i);
                Three f{};
                for (int a = 0; a < 3; ++a)
                    f[a] = obs * (p[(a + 1) % 3] + p[(a + 2) % 3]) + tr * p[(a + 1) % 3] * p[(a + 2) % 3];
                for (int dir = 0; dir < 4; ++dir) {
                    if ((dir == 2 && y == 0) || (dir == 3 && y == ny - 1)) continue;
                    std::size_t j;
                    if (dir == 0) j = static_cast<std::size_t>(y) * nx + (x == 0 ? nx - 1 : x - 1);
                    else if (dir == 1) j = static_cast<std::size_t>(y) * nx + (x == nx - 1 ? 0 : x + 1);
                    else if (dir == 2) j = i - nx;
                    else j = i + nx;
                    const Three q = F.cell(j);
                    const double dist = dir < 2 ? c.dx : c.dy;
                    for (int a = 0; a < 3; ++a)
                        for (int b = a + 1; b < 3; ++b) {
                            const double w = 2 * eps * sigma * (p[a] * q[b] - q[a] * p[b]) / (dist * dist);
                            f[a] += w * q[b];
                            f[b] -= w * q[a];
                        }
                }
                const double mean = (f[0] + f[1] + f[2]) / 3;
                for (int a = 0; a < 3; ++a) {
                    force[a * n + i] = f[a] - mean;
                    fmax = std::max(fmax, std::abs(force[a * n + i]));
                }
            }
        if (!std::isfinite(fmax)) throw std::runtime_error("nonfinite smoothing force");
        if (fmax == 0) {
            rep.steps = step - 1;
            rep.converged = true;
            rep.rel = 0;
            rep.change = 0;
            break;
        }
        double scale = std::min(base, c.smooth_max_phase_change / fmax);
        bool accepted = false;
        double e_new = previous, change = 0;
        for (int bt = 0; bt < 24 && !accepted; ++bt) {
            change = 0;
            for (std::size_t i = 0; i < n; ++i) {
                const Three p = F.cell(i);
                Three z{};
                for (int a = 0; a < 3; ++a) z[a] = p[a] - scale * force[a * n + i];
                z = project_simplex(z);
                for (int a = 0; a < 3; ++a) change = std::max(change, std::abs(z[a] - p[a]));
                trial.set(i, z);
            }
            e_new = interface_energy(trial, c);
            const double etol = 1e-14 * std::max({1.0, std::abs(previous), std::abs(rep.e0)});
            if (std::isfinite(e_new) && e_new <= previous + etol) accepted = true;
            else scale *= .5;
        }
        if (!accepted) throw std::runtime_error("sharp smoothing could not find an energy-decreasing step");
        F.phi = trial.phi;
        const double rel = std::abs(previous - e_new) / std::max(std::abs(previous), 1e-300);
        rep.steps = step;
        rep.e1 = e_new;
        rep.rel = rel;
        rep.change = change;
        previous = e_new;
        if (step >= c.smooth_min_steps && rel < c.smooth_energy_rtol && change < 1e-5) {
            rep.converged = true;
            break;
        }
    }
    if (rep.steps == 0) rep.e1 = rep.e0;
    rep.after = phase_fractions(F);
    return rep;
}

// ---------------------------------------------------------------------------
// HDF5 / XMF (pearl_serial.cpp:1854-2115)
// ---------------------------------------------------------------------------
void hcheck(herr_t s, const char* what) {
    if (s < 0) throw std::runtime_error(std::string("HDF5: ") + what);
}

struct H5 {
    hid_t id;
    herr_t (*close)(hid_t);
    H5(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {
        if (i < 0) throw std::runtime_error("HDF5 open/create failure");
    }
    ~H5() {
        if (id >= 0) close(id);
    }
    H5(const H5&) = delete;
    operator hid_t() const { return id; }
};

void attr_double(hid_t f, const char* name, double v) {
    H5 sp(H5Screate(H5S_SCALAR), H5Sclose);
    H5 a(H5Acreate2(f, name, H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    hcheck(H5Awrite(a, H5T_NATIVE_DOUBLE, &v), "write scalar");
}

double read_double(hid_t f, const char* name) {
    H5 a(H5Aopen(f, name, H5P_DEFAULT), H5Aclose);
    double v;
    hcheck(H5Aread(a, H5T_NATIVE_DOUBLE, &v), "read scalar");
    return v;
}

void attr_string(hid_t f, const char* name, const std::string& v) {
    H5 t(H5Tcopy(H5T_C_S1), H5Tclose);
    hcheck(H5Tset_size(t, v.size() + 1), "string size");
    H5 sp(H5Screate(H5S_SCALAR), H5Sclose);
    H5 a(H5Acreate2(f, name, t, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    hcheck(H5Awrite(a, t, v.c_str()), "string");
}

std::string read_string(hid_t f, const char* name) {
    H5 a(H5Aopen(f, name, H5P_DEFAULT), H5Aclose);
    H5 t(H5Aget_type(a), H5Tclose);
    std::vector<char> s(H5Tget_size(t) + 1, 0);
    hcheck(H5Aread(a, t, s.data()), "read string");
    return s.data();
}

void write_dataset(hid_t f, const std::string& name, const double* v, int nx, int ny) {
    hsize_t dims[2]{static_cast<hsize_t>(ny), static_cast<hsize_t>(nx)};
    hsize_t chunk[2]{static_cast<hsize_t>(std::min(ny, 64)), static_cast<hsize_t>(std::min(nx, 64))};
    H5 sp(H5Screate_simple(2, dims, nullptr), H5Sclose), prop(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
    hcheck(H5Pset_chunk(prop, 2, chunk), "chunk");
    hcheck(H5Pset_fletcher32(prop), "checksum");
    H5 set(H5Dcreate2(f, name.c_str(), H5T_IEEE_F64LE, sp, H5P_DEFAULT, prop, H5P_DEFAULT), H5Dclose);
    hcheck(H5Dwrite(set, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v), "dataset write");
}

std::vector<double> read_dataset(hid_t f, const char* name, int nx, int ny) {
    H5 set(H5Dopen2(f, name, H5P_DEFAULT), H5Dclose), sp(H5Dget_space(set), H5Sclose);
    hsize_t dims[2]{};
    if (H5Sget_simple_extent_ndims(sp) != 2) throw std::runtime_error("dataset is not 2D");
    H5Sget_simple_extent_dims(sp, dims, nullptr);
    if (dims[0] != static_cast<hsize_t>(ny) || dims[1] != static_cast<hsize_t>(nx))
        throw std::runtime_error("checkpoint dimensions mismatch");
    std::vector<double> v(static_cast<std::size_t>(nx) * ny);
    hcheck(H5Dread(set, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()), "dataset read");
    return v;
}

void write_xmf(const std::filesystem::path& h5, const std::string& grid, bool with_time, double time,
               double origin_y, int nx, int ny, double dx, double dy, const std::vector<const char*>& names,
               bool atomic) {
    auto xmf = h5;
    xmf.replace_extension(".xmf");
    auto tmp = xmf;
    if (atomic) tmp += ".part";
    std::ofstream out(tmp);
    if (!out) throw std::runtime_error("cannot write XMF");
    out << std::setprecision(17) << "<?xml version=\"1.0\"?>\n<Xdmf Version=\"3.0\"><Domain><Grid Name=\"" << grid
        << "\" GridType=\"Uniform\">\n";
    if (with_time) out << "<Time Value=\"" << time << "\"/>\n";
    out << "<Topology TopologyType=\"2DCoRectMesh\" Dimensions=\"" << ny << " " << nx << "\"/>\n"
        << "<Geometry GeometryType=\"ORIGIN_DXDY\"><DataItem Dimensions=\"2\" Format=\"XML\">" << origin_y << " "
        << .5 * dx << "</DataItem><DataItem Dimensions=\"2\" Format=\"XML\">" << dy << " " << dx
        << "</DataItem></Geometry>\n";
    for (const char* name : names)
        out << "<Attribute Name=\"" << name << "\" AttributeType=\"Scalar\" Center=\"Node\"><DataItem Dimensions=\""
            << ny << " " << nx << "\" NumberType=\"Float\" Precision=\"8\" Format=\"HDF\">" << h5.filename().string()
            << ":/" << name << "</DataItem></Attribute>\n";
    out << "</Grid></Domain></Xdmf>\n";
    out.close();
    if (!out) throw std::runtime_error(atomic ? "XMF write failure" : "initial XMF write failure");
    if (atomic) std::filesystem::rename(tmp, xmf);
}

struct Meta {
    double time = 0, dt = 0, shift_distance = 0, initial_carbon = 0, carbon_added = 0, carbon_removed = 0;
    long long step = 0, shifts = 0, rejected = 0;
};

// write_frame :1914-1968
void write_frame(const Fields& F, const Meta& s, const Config& c, double mu_ref, const std::filesystem::path& path) {
    if (std::filesystem::exists(path)) throw std::runtime_error("refuse overwriting frame " + path.string());
    auto tmp = path;
    tmp += ".part";
    {
        H5 f(H5Fcreate(tmp.c_str(), H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
        const char* names[3] = {"alpha", "theta", "gamma"};
        for (int a = 0; a < 3; ++a) write_dataset(f, names[a], F.phi.data() + a * F.n(), F.nx, F.ny);
        write_dataset(f, "carbon", F.c.data(), F.nx, F.ny);
        write_dataset(f, "mu", F.mu.data(), F.nx, F.ny);
        attr_string(f, "physics_signature", c.physics_signature());
        attr_string(f, "format", "pearlite-v0.4-5fields");
        attr_string(f, "thermo_mode", c.thermo_mode);
        attr_double(f, "temperature_K", c.T);
        if (c.thermo_mode == "eutectoid_reference") {
            attr_double(f, "T_eutectoid_K", c.T_eutectoid);
            attr_double(f, "undercooling_K", c.undercooling);
        }
        attr_double(f, "time", s.time);
        attr_double(f, "dt_next", s.dt);
        attr_double(f, "step", static_cast<double>(s.step));
        attr_double(f, "dx", c.dx);
        attr_double(f, "dy", c.dy);
        attr_double(f, "nx", c.nx);
        attr_double(f, "ny", c.ny);
        attr_double(f, "shift_distance", s.shift_distance);
        attr_double(f, "shifts", static_cast<double>(s.shifts));
        attr_double(f, "initial_carbon", s.initial_carbon);
        attr_double(f, "carbon_added", s.carbon_added);
        attr_double(f, "carbon_removed", s.carbon_removed);
        attr_double(f, "mu_reference", mu_ref);
        attr_double(f, "rejected_steps", static_cast<double>(s.rejected));
        hcheck(H5Fflush(f, H5F_SCOPE_GLOBAL), "flush");
    }
    std::filesystem::rename(tmp, path);
    write_xmf(path, "pearlite", true, s.time, .5 * c.dy + s.shift_distance, c.nx, c.ny, c.dx, c.dy,
              {"alpha", "theta", "gamma", "carbon", "mu"}, true);
}

// write_initial_phases :2075-2115
void write_initial_phases(const Fields& F, const Config& c, const std::filesystem::path& path, const std::string& stage,
                          int pseudo_step) {
    if (std::filesystem::exists(path)) throw std::runtime_error("initial snapshot exists");
    auto tmp = path;
    tmp += ".part";
    {
        H5 f(H5Fcreate(tmp.c_str(), H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
        const char* names[3] = {"alpha", "theta", "gamma"};
        for (int a = 0; a < 3; ++a) write_dataset(f, names[a], F.phi.data() + a * F.n(), F.nx, F.ny);
        attr_string(f, "format", "pearlite-preprocessing-phases-only");
        attr_string(f, "stage", stage);
        attr_double(f, "pseudo_step", pseudo_step);
        attr_double(f, "dx", c.dx);
        attr_double(f, "dy", c.dy);
        hcheck(H5Fflush(f, H5F_SCOPE_GLOBAL), "flush initial snapshot");
    }
    std::filesystem::rename(tmp, path);
    write_xmf(path, stage, false, 0, .5 * c.dy, c.nx, c.ny, c.dx, c.dy, {"alpha", "theta", "gamma"}, false);
}

bool valid_state(const Fields& F, const Thermo& f, bool phase_only) {  // :1096-1116
    bool good = true;
    for (std::size_t i = 0; i < F.n(); ++i) {
        double sum = 0;
        const Three p = F.cell(i);
        for (double a : p) {
            good = good && std::isfinite(a) && a >= -1e-12 && a <= 1 + 1e-12;
            sum += a;
        }
        good = good && std::abs(sum - 1) < 1e-10;
        good = good && std::isfinite(F.mu[i]) && std::isfinite(F.c[i]);
        if (!phase_only) {
            good = good && F.c[i] >= -1e-10 && F.c[i] <= 1 + 1e-10;
            for (int a = 0; a < 3; ++a)
                if (p[a] > .01) {
                    const double cp = f[a].c_at_mu(F.mu[i]);
                    good = good && cp >= -1e-10 && cp <= 1 + 1e-10;
                }
        }
    }
    return good;
}

double total_carbon(const Fields& F, double dx, double dy) {  // :784-788
    double v = 0;
    for (double c : F.c) v += c;
    return v * dx * dy;
}

// read_checkpoint :1970-2020
double read_checkpoint(Fields& F, Meta& s, const Config& c, const std::string& path) {
    std::array<double, 10> meta{};
    {
        H5 f(H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
        if (read_string(f, "format") != "pearlite-v0.4-5fields" || read_string(f, "physics_signature") != c.physics_signature())
            throw std::runtime_error("checkpoint physics signature mismatch");
        auto a = read_dataset(f, "alpha", c.nx, c.ny), t = read_dataset(f, "theta", c.nx, c.ny),
             g = read_dataset(f, "gamma", c.nx, c.ny);
        F.c = read_dataset(f, "carbon", c.nx, c.ny);
        F.mu = read_dataset(f, "mu", c.nx, c.ny);
        F.phi.resize(3 * F.n());
        std::copy(a.begin(), a.end(), F.phi.begin());
        std::copy(t.begin(), t.end(), F.phi.begin() + F.n());
        std::copy(g.begin(), g.end(), F.phi.begin() + 2 * F.n());
        meta = {read_double(f, "time"),         read_double(f, "dt_next"),        read_double(f, "step"),
                read_double(f, "shift_distance"), read_double(f, "shifts"),       read_double(f, "initial_carbon"),
                read_double(f, "carbon_added"),  read_double(f, "carbon_removed"), read_double(f, "mu_reference"),
                read_double(f, "rejected_steps")};
        for (double v : meta)
            if (!std::isfinite(v)) throw std::runtime_error("nonfinite checkpoint metadata");
        if (meta[0] < 0 || meta[1] <= 0 || meta[2] < 0 || meta[2] != std::floor(meta[2]) || meta[2] > 9e15)
            throw std::runtime_error("bad checkpoint time/step");
    }
    s.time = meta[0];
    s.dt = std::min(c.ctl.dt_max, meta[1]);
    s.step = static_cast<long long>(meta[2]);
    s.shift_distance = meta[3];
    s.shifts = static_cast<long long>(meta[4]);
    s.initial_carbon = meta[5];
    s.carbon_added = meta[6];
    s.carbon_removed = meta[7];
    s.rejected = static_cast<long long>(meta[9]);
    if (!valid_state(F, c.thermo, c.ctl.phase_only)) throw std::runtime_error("invalid checkpoint state");
    double closure = 0;
    for (std::size_t i = 0; i < F.n(); ++i)
        closure = std::max(closure, std::abs(F.c[i] - std::fma(mixture(F.cell(i), c.thermo).s, F.mu[i],
                                                               mixture(F.cell(i), c.thermo).c0)));
    if (closure > c.ctl.closure_tol) throw std::runtime_error("checkpoint closure failed");
    const double ledger = total_carbon(F, c.dx, c.dy) + s.carbon_removed - s.carbon_added - s.initial_carbon;
    if (!c.ctl.phase_only && std::abs(ledger) > 1e-10 * c.nx * c.ny * c.dx * c.dy)
        throw std::runtime_error("checkpoint mass ledger failed");
    return meta[8];
}

// write_thermodynamic_report :2118-2160
void write_thermo_report(const Config& c, const std::filesystem::path& path) {
    std::ofstream out(path);
    if (!out) throw std::runtime_error("cannot write thermodynamics report");
    out << std::setprecision(17);
    out << "thermo_mode=" << c.thermo_mode << "\nrun_temperature_K=" << c.T << "\nVm_m3_per_mol=" << c.Vm << "\n";
    out << "energy_convention=legacy v0.3 molar-code convention; chemical term Psi/Vm, c=atomic fraction\n";
    out << "unit_scope=this preserves the inherited conversion; it does not certify the material's physical molar volume\n";
    out << "thermal_history=isothermal at the run temperature; no finite-rate cooling segment\n";
    if (c.thermo_mode == "eutectoid_reference") {
        out << "T_eutectoid_K=" << c.T_eutectoid << "\nundercooling_K=" << c.undercooling
            << "\nmax_undercooling_K=" << c.max_undercooling << "\ntemperature_law=" << c.temperature_law << "\n";
        if (c.temperature_law == "phase_diagram_fixed_curvature") {
            out << "phase_diagram_slope_units=carbon mole fraction/K (dc_eq/dT)\n";
            out << "phase_diagram_curvature=fixed at supplied reference Aeq\n";
            out << "phase_diagram_gamma_reference=ABC held at reference values\n";
            out << "phase_diagram_reference_c_alpha_theta_gamma=" << c.diagram.ceq[0] << ',' << c.diagram.ceq[1] << ','
                << c.diagram.ceq[2] << "\n";
            out << "phase_diagram_slopes_alpha_ag_gamma_ag_theta_tg_gamma_tg=";
            for (int j = 0; j < 4; ++j) out << (j ? "," : "") << c.diagram.slope[j];
            out << "\nphase_diagram_reference_residual=preserved, not retuned\n";
        }
    }
    const std::array<std::string, 3> names{"alpha", "theta", "gamma"};
    for (int i = 0; i < 3; ++i) {
        const auto& n = names[i];
        if (c.thermo_mode == "eutectoid_reference") {
            const auto& p = c.reference_thermo[i];
            const auto& sl = c.slope[i];
            out << "reference_" << n << "_ABC=" << p.A << ',' << p.B << ',' << p.C << "\n";
            if (c.temperature_law == "linear_gp") out << "dq_dT_" << n << "_q2_q1_q0=" << sl.q2 << ',' << sl.q1 << ',' << sl.q0 << "\n";
        }
        const auto& p = c.thermo[i];
        const GP g = to_gp(p);
        out << "effective_" << n << "_ABC=" << p.A << ',' << p.B << ',' << p.C << "\n";
        out << "effective_" << n << "_q2_q1_q0=" << g.q2 << ',' << g.q1 << ',' << g.q0 << "\n";
        out << "effective_" << n << "_chi=" << p.chi() << "\nD_" << n << "=" << p.D << "\n";
    }
    out << "initial_composition_mode=" << c.composition_mode << "\n";
    if (c.composition_mode != "mean") out << "initial_mu_at_run_temperature=" << c.initial_mu << "\n";
    else out << "initial_mu_at_run_temperature=solved after smoothing from specified mean_carbon\n";
    out << "initialization_note=one local mu field; tip samples at distinct positions need not be equal\n";
    out << "tau_at_ag_tg=" << c.tau[0] << ',' << c.tau[1] << ',' << c.tau[2] << "\n";
    out << "tau_note=explicit input, not recalibrated by changing temperature; physical applicability remains a separate check\n";
    out.close();
    if (!out) throw std::runtime_error("thermodynamics report write failed");
}

// initialize :1761-1837
using Observer = std::function<void(const Fields&, const char*, int)>;

void label_cells(Fields& F, const Config& c) {  // the labelling loop :1764-1795
    const int nx = c.nx, ny = c.ny;
    const double width = c.theta_columns * c.dx, L = nx * c.dx, eps = c.eps;
    for (int y = 1; y <= ny; ++y)
        for (int x = 1; x <= nx; ++x) {
            const double xx = (x - .5) * c.dx, yy = (y - .5) * c.dy;
            Three p{};
            if (c.initialization == "lamella") {
                const double dist = xx < width ? -std::min(xx, width - xx) : std::min(xx - width, L - xx);
                const double th = profile(dist, eps), product = profile(yy - c.seed_height, eps);
                p = {(1 - th) * product, th * product, 1 - product};
            } else if (c.initialization == "microsim") {
                p = fill_phase_at(c.fills, x - 1, y - 1);
            } else if (c.initialization == "sharp_smooth") {
                const int gxi = x - 1;
                const bool product = yy < c.seed_height;
                if (!product) p = {0, 0, 1};
                else if (gxi < c.theta_columns) p = {0, 1, 0};
                else p = {1, 0, 0};
            } else {
                const double dist = c.initialization == "flat" ? yy - c.seed_height
                                                                : std::hypot(xx - .5 * L, yy - .5 * ny * c.dy) - c.radius;
                const double u = profile(dist, eps);
                p[c.phase_a] = u;
                p[c.phase_b] = 1 - u;
            }
            F.set(static_cast<std::size_t>(y - 1) * nx + (x - 1), p);
        }
}

bool sharp_labels(const Config& c) { return c.initialization == "sharp_smooth" || c.initialization == "microsim"; }

double initialize(Fields& F, Meta& s, const Config& c, SmoothReport* rep, const Observer& obs) {
    const int nx = c.nx, ny = c.ny;
    F.nx = nx;
    F.ny = ny;
    F.phi.assign(3 * F.n(), 0.0);
    F.c.assign(F.n(), 0.0);
    F.mu.assign(F.n(), 0.0);
    label_cells(F, c);
    if (sharp_labels(c)) {
        if (obs) obs(F, "init_sharp", 0);
        SmoothReport r = smooth(F, c);
        if (rep) *rep = r;
        if (obs) obs(F, "init_smoothed", r.steps);
    } else if (rep) {
        rep->before = phase_fractions(F);
        rep->after = rep->before;
        rep->e0 = interface_energy(F, c);
        rep->e1 = rep->e0;
        rep->converged = true;
    }
    double mu = c.initial_mu;
    if (c.composition_mode == "mean") {
        double si = 0, sc = 0;
        for (std::size_t i = 0; i < F.n(); ++i) {
            const Mix m = mixture(F.cell(i), c.thermo);
            si += m.c0;
            sc += m.s;
        }
        mu = (nx * ny * c.initial_carbon - si) / sc;
    }
    for (std::size_t i = 0; i < F.n(); ++i) {
        const Mix m = mixture(F.cell(i), c.thermo);
        F.mu[i] = mu;
        F.c[i] = std::fma(m.s, mu, m.c0);
    }
    s = Meta{};
    s.dt = c.dt;
    s.initial_carbon = total_carbon(F, c.dx, c.dy);
    if (!valid_state(F, c.thermo, c.ctl.phase_only)) throw std::runtime_error("initial state out of model range");
    return mu;
}

// ---------------------------------------------------------------------------
// GPU context helpers
// ---------------------------------------------------------------------------
void check(pearl_ctx* g, int st) {
    if (st != PEARL_OK) throw std::runtime_error(g ? pearl_last_error(g) : "pearl_gpu call failed");
}

// The sharp labelling of initialize as pearl_fill operations: the FILL
// lines (microsim), or sharp_smooth's two boxes -- theta in the first
// theta_columns, alpha in the rest, of the rows whose centre (gy + 0.5) dy
// lies below seed_height (:1777-1786).  The directions are the reference's
// expressions (:1423-1424, :1440-1441, :1487-1488).
std::vector<pearl_fill> fill_ops(const Config& c) {
    static const char* const kMethods[] = {"FILLCUBE", "FILLCYLINDER", "FILLCENTERBOX2D", "FILLYJUNCTION2D",
                                           "FILLYJUNCTIONLAMELLAE2D", "FILLADHEREDPEARLITEARM2D",
                                           "FILLPEARLITEELLIPSE2D"};
    const double pi = 3.14159265358979323846;
    std::vector<pearl_fill> ops;
    if (c.initialization == "microsim") {
        for (const auto& f : c.fills) {
            pearl_fill o{};
            o.method = -1;
            for (int k = 0; k < 7; ++k)
                if (f.method == kMethods[k]) o.method = k;
            if (o.method < 0 || f.p.size() > PEARL_FILL_PARAMS)
                throw std::invalid_argument("unsupported filling method: " + f.method);
            std::copy(f.p.begin(), f.p.end(), o.p);
            auto direction = [&](int k, double degrees) {
                const double a = degrees * pi / 180.0;
                o.dir[k][0] = std::cos(a);
                o.dir[k][1] = std::sin(a);
            };
            if (o.method == PEARL_FILL_YJUNCTION2D || o.method == PEARL_FILL_YJUNCTIONLAMELLAE2D)
                for (int k = 0; k < 3; ++k) direction(k, f.p[6 + k]);
            else if (o.method == PEARL_FILL_ADHEREDPEARLITEARM2D || o.method == PEARL_FILL_PEARLITEELLIPSE2D)
                direction(0, f.p[2]);
            ops.push_back(o);
        }
    } else {
        int rows = 0;
        while (rows < c.ny && (rows + .5) * c.dy < c.seed_height) ++rows;
        const int boxes[2][3] = {{1, 0, c.theta_columns - 1}, {0, c.theta_columns, c.nx - 1}};
        for (const auto& b : boxes) {
            pearl_fill o{};
            o.method = PEARL_FILL_CUBE;
            o.p[0] = b[0];
            o.p[1] = b[1];
            o.p[4] = b[2];
            o.p[5] = rows - 1;
            ops.push_back(o);
        }
    }
    return ops;
}

// --gpu-init (no reference counterpart): initialize with its per-cell work
// on the GPU, on a context of its own.  Sharp labels (pearl_fill_phases) and
// the capillarity-only smoothing (pearl_smooth) for sharp_smooth and
// microsim; the diffuse profiles of the other initializations stay on the
// host.  Then the carbon set-up (pearl_init_composition).  The labels and
// the per-cell carbon are the reference's bit for bit; the smoothing's
// energies, a mean_carbon mu and the initial carbon are summed in the
// blocked order.
double initialize_device(Fields& F, Meta& s, const Config& c, SmoothReport* rep, const Observer& obs, int device) {
    if (sharp_labels(c) && (c.smooth_steps < 1 || c.smooth_min_steps < 0 || c.smooth_min_steps > c.smooth_steps ||
                            !(c.smooth_energy_rtol > 0) || !(c.smooth_max_phase_change > 0) ||
                            c.smooth_max_phase_change > .2))
        throw std::invalid_argument("invalid sharp-smoothing controls");
    F.nx = c.nx;
    F.ny = c.ny;
    F.phi.assign(3 * F.n(), 0.0);
    F.c.assign(F.n(), 0.0);
    F.mu.assign(F.n(), 0.0);
    pearl_ctx* g = nullptr;
    check(nullptr, pearl_create(&c.ctl, c.initial_mu, PEARL_FP64, device, 0, &g));
    struct Destroy {
        pearl_ctx* g;
        ~Destroy() { pearl_destroy(g); }
    } guard{g};
    if (sharp_labels(c)) {
        const std::vector<pearl_fill> ops = fill_ops(c);
        check(g, pearl_fill_phases(g, ops.data(), static_cast<int>(ops.size())));
        check(g, pearl_get_state(g, F.phi.data(), nullptr, nullptr, nullptr));
        if (obs) obs(F, "init_sharp", 0);
        SmoothReport r;
        r.before = phase_fractions(F);
        pearl_smooth_report sr{};
        check(g, pearl_smooth(g, c.smooth_steps, c.smooth_min_steps, c.smooth_energy_rtol, c.smooth_max_phase_change,
                              &sr));
        check(g, pearl_get_state(g, F.phi.data(), nullptr, nullptr, nullptr));
        r.steps = sr.steps;
        r.converged = sr.converged != 0;
        r.e0 = sr.energy_initial;
        r.e1 = sr.energy_final;
        r.rel = sr.final_relative_energy_change;
        r.change = sr.final_max_phase_change;
        r.after = phase_fractions(F);
        if (rep) *rep = r;
        if (obs) obs(F, "init_smoothed", r.steps);
    } else {
        label_cells(F, c);
        if (rep) {
            rep->before = phase_fractions(F);
            rep->after = rep->before;
            rep->e0 = interface_energy(F, c);
            rep->e1 = rep->e0;
            rep->converged = true;
        }
        pearl_scalars sc{};
        check(g, pearl_set_state(g, F.phi.data(), F.c.data(), F.mu.data(), &sc));
    }
    const bool mean = c.composition_mode == "mean";
    double mu = 0, carbon = 0;
    int valid = 0;
    check(g, pearl_init_composition(g, mean ? PEARL_COMPOSITION_MEAN : PEARL_COMPOSITION_MU,
                                    mean ? c.initial_carbon : c.initial_mu, &mu, &carbon, &valid));
    check(g, pearl_get_state(g, nullptr, F.c.data(), F.mu.data(), nullptr));
    s = Meta{};
    s.dt = c.dt;
    s.initial_carbon = carbon;
    if (!valid) throw std::runtime_error("initial state out of model range");
    return mu;
}

pearl_scalars to_scalars(const Meta& m) {
    return {m.time, m.dt, m.shift_distance, m.initial_carbon, m.carbon_added, m.carbon_removed, m.step, m.shifts, m.rejected};
}

Meta from_scalars(const pearl_scalars& s) {
    Meta m;
    m.time = s.time;
    m.dt = s.dt;
    m.shift_distance = s.shift_distance;
    m.initial_carbon = s.initial_carbon;
    m.carbon_added = s.carbon_added;
    m.carbon_removed = s.carbon_removed;
    m.step = s.step;
    m.shifts = s.shifts;
    m.rejected = s.rejected;
    return m;
}

}  // namespace host

// ---------------------------------------------------------------------------
// main :2184-2336
// ---------------------------------------------------------------------------
namespace {
volatile std::sig_atomic_t interrupted = 0;
void on_signal(int) { interrupted = 1; }
}  // namespace

int main(int argc, char** argv) {
    using namespace host;
    int status = 0;
    pearl_ctx* gpu = nullptr;
    try {
        if (argc < 3)
            throw std::invalid_argument(
                "Usage: pearl input.in NEW_OUTPUT_DIR [--restart frame.h5] [--px N --py N] [--max-steps N] [--thermo-only]");
        std::string input = argv[1], output = argv[2], restart;
        bool thermo_only = false;
        int px = 1, py = 1, device = 0;
        bool gpu_init = false;
        long long stop_steps = -1;
        for (int k = 3; k < argc; ++k) {
            const std::string key = argv[k];
            if (key == "--thermo-only") {
                thermo_only = true;
                continue;
            }
            if (key == "--gpu-init") {  // no reference counterpart: smoothing on the GPU
                gpu_init = true;
                continue;
            }
            if (k + 1 == argc) throw std::invalid_argument("missing CLI value");
            const std::string val = argv[++k];
            if (key == "--restart") restart = val;
            else if (key == "--px") px = std::stoi(val);
            else if (key == "--py") py = std::stoi(val);
            else if (key == "--max-steps") stop_steps = std::stoll(val);
            else if (key == "--device") device = std::stoi(val);
            else throw std::invalid_argument("unknown CLI flag: " + key);
        }
        if (thermo_only && !restart.empty()) throw std::invalid_argument("--thermo-only cannot be combined with --restart");
        Config config = read_config(input);
        if (stop_steps == 0 || stop_steps < -1) throw std::invalid_argument("max-steps must be positive");
        if (config.nx < 3 || config.ny < 2 || config.dx <= 0 || config.dy <= 0) throw std::invalid_argument("invalid grid");
        // no reference counterpart: the device context's size limit (pearl_params)
        if (3 * static_cast<long long>(config.nx) * config.ny >= (1LL << 31) - (1LL << 20))
            throw std::invalid_argument("grid too large for one device context (3*nx*ny must stay below 2^31 - 2^20)");
        if (px != 1 || py != 1) throw std::invalid_argument("serial build requires px=py=1");
        {
            bool ok = true;
            try {
                if (std::filesystem::exists(output)) throw std::runtime_error("output path already exists; refuse overwrite");
                std::filesystem::create_directories(output);
                std::filesystem::copy_file(input, std::filesystem::path(output) / "Input.in");
                write_thermo_report(config, std::filesystem::path(output) / "thermodynamics.txt");
            } catch (...) {
                ok = false;
            }
            if (!ok) throw std::runtime_error("cannot create NEW output directory");
        }
        if (thermo_only) {
            std::cout << "THERMO_ONLY: report written; no initialization or evolution performed\n";
            return 0;
        }
        Fields F;
        F.nx = config.nx;
        F.ny = config.ny;
        Meta state;
        SmoothReport rep;
        Observer writer = [&](const Fields& fl, const char* stage, int pseudo) {
            write_initial_phases(fl, config, std::filesystem::path(output) / (std::string(stage) + ".h5"), stage, pseudo);
        };
        const double mu_reference =
            restart.empty() ? (gpu_init ? initialize_device(F, state, config, &rep, writer, device)
                                        : initialize(F, state, config, &rep, writer))
                            : read_checkpoint(F, state, config, restart);
        if (restart.empty()) {
            std::ofstream log(std::filesystem::path(output) / "initialization_report.txt");
            log << std::setprecision(17);
            log << "initialization=" << config.initialization << "\n";
            log << "physical_time_after_preprocess=" << state.time << "\n";
            log << "solver_step_after_preprocess=" << state.step << "\n";
            log << "smooth_steps_completed=" << rep.steps << "\n";
            log << "smooth_energy_converged=" << (rep.converged ? 1 : 0) << "\n";
            log << "smooth_energy_initial=" << rep.e0 << "\n";
            log << "smooth_energy_final=" << rep.e1 << "\n";
            log << "smooth_final_relative_energy_change=" << rep.rel << "\n";
            log << "smooth_final_max_phase_change=" << rep.change << "\n";
            log << "phase_fraction_before=" << rep.before[0] << ',' << rep.before[1] << ',' << rep.before[2] << "\n";
            log << "phase_fraction_after=" << rep.after[0] << ',' << rep.after[1] << ',' << rep.after[2] << "\n";
            log << "mu_reference=" << mu_reference << "\n";
            log << "NOTE=smoothing is capillarity-only preprocessing pseudo-time; carbon diffusion is disabled and physical "
                   "time is reset to zero before coupled evolution.\n";
        }
        if (state.time >= config.end_time) throw std::runtime_error("checkpoint time >= end_time");
        // The state moves to the GPU once; the loop below runs there.
        check(nullptr, pearl_create(&config.ctl, mu_reference, PEARL_FP64, device, 0, &gpu));
        pearl_scalars sc = to_scalars(state);
        check(gpu, pearl_set_state(gpu, F.phi.data(), F.c.data(), F.mu.data(), &sc));
        std::signal(SIGINT, on_signal);
        std::signal(SIGTERM, on_signal);
        const long long first_step = state.step;
        long long frames = 0;
        double next_output = (std::floor((state.time + 1e-12 * config.output_dt) / config.output_dt) + 1.0) * config.output_dt;
        pearl_step_info last{};
        long long cg_total = 0;
        int cg_peak = 0;
        const auto start = std::chrono::steady_clock::now();
        std::ofstream log(std::filesystem::path(output) / "diagnostics.csv");
        if (!log) throw std::runtime_error("cannot open diagnostics");
        log << "time,step,dt_next,mean_carbon,mass_ledger_error,reduced_energy,phi_sum_error,c_min,c_max,mu_min,mu_max,"
               "cg_last,cg_peak,closure_last,rejected,shift_distance,wall_s\n";
        std::cout << "resources: ranks=1 layout=1x1 grid=" << config.nx << "x" << config.ny << " device=" << device << "\n";
        auto snapshot = [&]() {
            pearl_diag dg{};
            check(gpu, pearl_diagnostics(gpu, &dg));
            check(gpu, pearl_get_state(gpu, F.phi.data(), F.c.data(), F.mu.data(), &sc));
            state = from_scalars(sc);
            const double ledger = dg.mass + state.carbon_removed - state.carbon_added - state.initial_carbon;
            const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
            std::ostringstream fn;
            fn << "frame_" << std::setfill('0') << std::setw(6) << frames++ << ".h5";
            write_frame(F, state, config, mu_reference, std::filesystem::path(output) / fn.str());
            log << std::setprecision(17) << state.time << ',' << state.step << ',' << state.dt << ','
                << dg.mass / (config.nx * config.ny * config.dx * config.dy) << ',' << ledger << ',' << dg.energy << ','
                << dg.phi_sum_error << ',' << dg.c_min << ',' << dg.c_max << ',' << dg.mu_min << ',' << dg.mu_max << ','
                << last.cg_iterations << ',' << cg_peak << ',' << last.closure << ',' << state.rejected << ','
                << state.shift_distance << ',' << wall << '\n';
            log.flush();
            std::cout << "t=" << state.time << " step=" << state.step << " dt_next=" << state.dt << " mass_error=" << ledger
                      << " CG=" << last.cg_iterations << " rejected=" << state.rejected << "\n";
        };
        snapshot();
        bool stopped = false;
        double saved_time = state.time;
        // The steps between two outputs run on the device (pearl_run) and end
        // where the reference's loop takes the output (time >= next_output -
        // eps, :2307).  A call covers at most kRunChunk steps, so the
        // interrupt flag is still seen between calls.
        const double eps = 1e-12 * std::max(config.end_time, config.output_dt);
        check(gpu, pearl_set_stop_tolerance(gpu, eps));
        constexpr long long kRunChunk = 512;
        while (sc.time < config.end_time) {
            if (interrupted || (stop_steps > 0 && sc.step - first_step >= stop_steps)) {
                stopped = true;
                break;
            }
            if (sc.step >= config.max_steps) throw std::runtime_error("configured max_steps reached before end_time");
            long long budget = std::min<long long>(kRunChunk, config.max_steps - sc.step);
            if (stop_steps > 0) budget = std::min<long long>(budget, stop_steps - (sc.step - first_step));
            long long taken = 0;
            check(gpu, pearl_run(gpu, std::min(config.end_time, next_output), budget, &taken, &cg_total, &cg_peak,
                                 &last));
            check(gpu, pearl_get_state(gpu, nullptr, nullptr, nullptr, &sc));
            if (sc.time >= next_output - eps || sc.time >= config.end_time - eps) {
                // Snap only the clock within rounding tolerance (:2308-2310).
                bool snapped = false;
                if (std::abs(sc.time - next_output) <= eps) {
                    sc.time = next_output;
                    snapped = true;
                }
                if (std::abs(sc.time - config.end_time) <= eps) {
                    sc.time = config.end_time;
                    snapped = true;
                }
                if (snapped) check(gpu, pearl_set_state(gpu, nullptr, nullptr, nullptr, &sc));
                int shifted = 0;
                double front = -1;
                check(gpu, pearl_maybe_shift(gpu, config.moving_window, config.shift_trigger, config.shift_target,
                                             config.shift_gamma_carbon, &shifted, &front));
                snapshot();
                saved_time = sc.time;
                next_output += config.output_dt;
            }
        }
        if (sc.time != saved_time) snapshot();
        check(gpu, pearl_get_state(gpu, F.phi.data(), F.c.data(), F.mu.data(), &sc));
        state = from_scalars(sc);
        write_frame(F, state, config, mu_reference, std::filesystem::path(output) / "checkpoint.h5");
        std::ofstream fin(std::filesystem::path(output) / (stopped ? "STOPPED.txt" : "COMPLETED.txt"));
        fin << std::setprecision(17) << "status=" << (stopped ? "USER_STOPPED" : "HORIZON_COMPLETED") << "\ntime="
            << state.time << "\nstep=" << state.step << "\nCG_total=" << cg_total << "\nrejected=" << state.rejected
            << "\nThis is a solver status, not certification of cooperative growth.\n";
    } catch (const std::exception& e) {
        std::cerr << "ERROR: " << e.what() << '\n';
        status = 1;
    }
    if (gpu) pearl_destroy(gpu);
    return status;
}
