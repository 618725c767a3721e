// Admission rules of the FILL operations of initialization = microsim (the
// reference's validate_filling_operations, pearl_serial.cpp:1251-1390),
// kept as data: each method has its parameter count and an ordered list of
// checks over its parameters; the first failing check names the error
// (the reference's texts, prefixed by the method).  The labelling itself
// runs on the device (pearl_fill_phases); this file only decides which
// operations are admitted and converts them to pearl_fill records.
#pragma once

#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "input_file.hpp"
#include "pearl_gpu.h"

namespace pearl_host {

namespace fill_rules {

// Check kinds.  Indices are 0-based parameter positions.
enum Op {
    INT,        // a: representable integer
    PHASE,      // a: integer phase ID 0..2
    DISTINCT,   // a, b and c (unless c < 0): phase IDs, pairwise different
    POS,        // a: > 0
    NONNEG,     // a: >= 0
    FRACTION,   // a: in (0, 1)
    ANGLE,      // a degrees convert to finite radians
    SQUARE,     // a^2 finite and not underflowing to 0
    ZERO,       // a and b equal 0 (the 2D z bounds), message `text`
    BOX,        // parameters 0..3 integers, 0 <= xlo < xhi <= nx, 0 <= ylo < yhi <= ny
    CUBE_BOX,   // inclusive cube bounds 1,2 .. 4,5 inside the grid
    CENTER,     // a, b inside the grid; integers when c != 0
    PAIRS_MAX,  // a <= INT_MAX / 2
    NOT_ABOVE,  // p[a] <= p[b], message `text`
    DERIVED,    // a length derived from the parameters (see derived()) finite and positive, named `text`
    REACH,      // the ellipse's endpoint distance squared is finite
};

struct Check {
    Op op;
    int a = 0, b = 0, c = 0;
    const char* text = "";
};

struct Method {
    const char* name;
    int method_code;  // PEARL_FILL_* (-1: refused)
    std::size_t count;
    std::vector<Check> checks;
    const char* refusal = "";
};

// Lengths the arm and ellipse geometries derive from their parameters.
enum Derived { LAMELLA_PERIOD, ARM_PERIOD, ARM_FERRITE, ARM_CEMENTITE, ELLIPSE_LENGTH, ELLIPSE_MAJOR, ELLIPSE_MINOR,
               ELLIPSE_CEMENTITE };

inline double derived(int which, const std::vector<double>& p) {
    switch (which) {
        case LAMELLA_PERIOD: return p[11] + p[12];
        case ARM_PERIOD: return p[3] / p[4];
        case ARM_FERRITE: return (1.0 - p[5]) * (p[3] / p[4]);
        case ARM_CEMENTITE: return p[5] * (p[3] / p[4]);
        case ELLIPSE_LENGTH: return p[4] * p[5];
        case ELLIPSE_MAJOR: return 0.5 * (p[4] * p[5]);
        case ELLIPSE_MINOR: return p[6] * (0.5 * (p[4] * p[5]));
        default: return p[7] * p[5];  // ELLIPSE_CEMENTITE
    }
}

inline const std::vector<Method>& methods() {
    static const std::vector<Method> table = {
        {"FILLCUBE", PEARL_FILL_CUBE, 7,
         {{PHASE, 0}, {INT, 1}, {INT, 2}, {INT, 3}, {INT, 4}, {INT, 5}, {INT, 6},
          {ZERO, 3, 6, 0, "2D filling requires zlo=zhi=0"}, {CUBE_BOX}}},
        {"FILLCENTERBOX2D", PEARL_FILL_CENTERBOX2D, 6, {{BOX}, {DISTINCT, 4, 5, -1}}},
        {"FILLYJUNCTION2D", PEARL_FILL_YJUNCTION2D, 13,
         {{BOX}, {CENTER, 4, 5, 1}, {POS, 9}, {POS, 10}, {ANGLE, 6}, {ANGLE, 7}, {ANGLE, 8},
          {DISTINCT, 11, 12, -1}}},
        {"FILLYJUNCTIONLAMELLAE2D", PEARL_FILL_YJUNCTIONLAMELLAE2D, 16,
         {{BOX}, {CENTER, 4, 5, 1}, {POS, 9}, {POS, 10}, {ANGLE, 6}, {ANGLE, 7}, {ANGLE, 8}, {POS, 11}, {POS, 12},
          {DISTINCT, 13, 14, 15}, {DERIVED, LAMELLA_PERIOD, 0, 0, "lamellar period"}}},
        {"FILLCYLINDER", PEARL_FILL_CYLINDER, 6,
         {{PHASE, 0}, {CENTER, 1, 2, 1}, {INT, 3}, {INT, 4}, {POS, 5},
          {ZERO, 3, 4, 0, "2D cylinder cross-section requires zlo=zhi=0"}, {SQUARE, 5}}},
        {"FILLADHEREDPEARLITEARM2D", PEARL_FILL_ADHEREDPEARLITEARM2D, 11,
         {{CENTER, 0, 1, 0}, {POS, 3}, {INT, 4}, {POS, 4}, {PAIRS_MAX, 4}, {FRACTION, 5}, {POS, 6}, {NONNEG, 7},
          {NONNEG, 8}, {DISTINCT, 9, 10, -1},
          {NOT_ABOVE, 7, 6, 0, "contact_half_width must not exceed peak_half_width"}, {ANGLE, 2}, {SQUARE, 3},
          {SQUARE, 6}, {SQUARE, 8}, {DERIVED, ARM_PERIOD, 0, 0, "arm period"},
          {DERIVED, ARM_FERRITE, 0, 0, "ferrite segment length"},
          {DERIVED, ARM_CEMENTITE, 0, 0, "cementite segment length"}}},
        {"FILLPEARLITEELLIPSE2D", PEARL_FILL_PEARLITEELLIPSE2D, 11,
         {{CENTER, 0, 1, 0}, {NONNEG, 3}, {INT, 4}, {POS, 4}, {POS, 5}, {POS, 6}, {FRACTION, 7}, {NONNEG, 8},
          {DISTINCT, 9, 10, -1}, {ANGLE, 2}, {SQUARE, 8}, {DERIVED, ELLIPSE_LENGTH, 0, 0, "ellipse length"},
          {DERIVED, ELLIPSE_MAJOR, 0, 0, "ellipse semimajor axis"},
          {DERIVED, ELLIPSE_MINOR, 0, 0, "ellipse semiminor axis"},
          {DERIVED, ELLIPSE_CEMENTITE, 0, 0, "cementite lamellar width"}, {REACH}}},
        {"FILLTHREEGRAINYJUNCTION2D", -1, 0, {},
         "historical multi-grain initialization requires separate gamma grain fields; "
         "Pearl has one gamma field and will not merge grain IDs silently"},
        {"FILLYJUNCTIONPEARLITE2D", -1, 0, {},
         "historical multi-grain initialization requires separate gamma grain fields; "
         "Pearl has one gamma field and will not merge grain IDs silently"},
    };
    return table;
}

// The first failing check's text, or an empty string.
inline std::string first_failure(const Method& m, const std::vector<double>& p, int nx, int ny) {
    auto nth = [](int i) { return "parameter " + std::to_string(i + 1); };
    auto is_int = [&](int i) {
        return std::trunc(p[i]) == p[i] && p[i] >= static_cast<double>(std::numeric_limits<int>::min()) &&
               p[i] <= static_cast<double>(std::numeric_limits<int>::max());
    };
    const std::string not_int = " must be a representable integer";
    const std::string bad_phase = "phase IDs must be alpha=0, theta=1 or gamma=2; gamma grain clones are unsupported";
    if (p.size() != m.count) return "expected " + std::to_string(m.count) + " parameters";
    for (double v : p)
        if (!std::isfinite(v)) return "parameters must be finite";
    for (const Check& k : m.checks) {
        switch (k.op) {
            case INT:
                if (!is_int(k.a)) return nth(k.a) + not_int;
                break;
            case PHASE:
                if (!is_int(k.a)) return nth(k.a) + not_int;
                if (p[k.a] < 0 || p[k.a] > 2) return bad_phase;
                break;
            case DISTINCT: {
                const int ids[3] = {k.a, k.b, k.c};
                const int n = k.c < 0 ? 2 : 3;
                for (int i = 0; i < n; ++i) {
                    if (!is_int(ids[i])) return nth(ids[i]) + not_int;
                    if (p[ids[i]] < 0 || p[ids[i]] > 2) return bad_phase;
                }
                for (int i = 0; i < n; ++i)
                    for (int j = i + 1; j < n; ++j)
                        if (p[ids[i]] == p[ids[j]]) return "listed phase IDs must be distinct";
                break;
            }
            case POS:
                if (!(p[k.a] > 0.0)) return nth(k.a) + " must be positive";
                break;
            case NONNEG:
                if (p[k.a] < 0.0) return nth(k.a) + " must be nonnegative";
                break;
            case FRACTION:
                if (!(p[k.a] > 0.0 && p[k.a] < 1.0)) return "cementite fraction must be strictly between 0 and 1";
                break;
            case ANGLE:
                if (!std::isfinite(p[k.a] * 3.14159265358979323846 / 180.0)) return "angle conversion to radians overflows";
                break;
            case SQUARE: {
                const double s = p[k.a] * p[k.a];
                if (!std::isfinite(s) || (p[k.a] > 0.0 && s == 0.0)) return "length squared overflows or underflows";
                break;
            }
            case ZERO:
                if (p[k.a] != 0 || p[k.b] != 0) return k.text;
                break;
            case BOX:
                for (int i = 0; i < 4; ++i)
                    if (!is_int(i)) return nth(i) + not_int;
                if (!(0 <= p[0] && p[0] < p[1] && p[1] <= nx && 0 <= p[2] && p[2] < p[3] && p[3] <= ny))
                    return "box must satisfy 0<=xlo<xhi<=nx and 0<=ylo<yhi<=ny";
                break;
            case CUBE_BOX:
                if (!(0 <= p[1] && p[1] <= p[4] && p[4] < nx && 0 <= p[2] && p[2] <= p[5] && p[5] < ny))
                    return "inclusive bounds must satisfy 0<=xlo<=xhi<nx and 0<=ylo<=yhi<ny";
                break;
            case CENTER:
                if (k.c && !is_int(k.a)) return nth(k.a) + not_int;
                if (k.c && !is_int(k.b)) return nth(k.b) + not_int;
                if (!(0 <= p[k.a] && p[k.a] < nx && 0 <= p[k.b] && p[k.b] < ny))
                    return "center/junction must be inside the physical grid";
                break;
            case PAIRS_MAX:
                if (p[k.a] > std::numeric_limits<int>::max() / 2) return "pair_count is too large";
                break;
            case NOT_ABOVE:
                if (p[k.a] > p[k.b]) return k.text;
                break;
            case DERIVED: {
                const double v = derived(k.a, p);
                if (!(v > 0.0) || !std::isfinite(v)) return std::string(k.text) + " must remain finite and positive";
                break;
            }
            case REACH: {
                const double reach = p[3] + p[4] * p[5];
                if (!std::isfinite(reach * reach)) return "ellipse endpoint distance overflows";
                break;
            }
        }
    }
    return {};
}

}  // namespace fill_rules

// Admit the FILL lines of an input (every error names its method) and
// return them as pearl_fill records: parameters in input order and the
// cosine and sine of each direction the operation uses, with the
// reference's expression cos(angle*pi/180.0) and its literal pi (:1396).
inline std::vector<pearl_fill> admit_fills(const std::vector<FillLine>& lines, int nx, int ny) {
    if (nx <= 0 || ny <= 0) throw std::invalid_argument("Filling requires positive nx and ny");
    std::vector<pearl_fill> out;
    for (const FillLine& f : lines) {
        const fill_rules::Method* m = nullptr;
        for (const auto& cand : fill_rules::methods())
            if (f.method == cand.name) m = &cand;
        if (!m) throw std::invalid_argument(f.method +
                                            ": unsupported filling method in the 2D three-phase Pearl solver; "
                                            "see FILLING_METHODS.md");
        if (m->method_code < 0) throw std::invalid_argument(f.method + ": " + m->refusal);
        const std::string why = fill_rules::first_failure(*m, f.p, nx, ny);
        if (!why.empty()) throw std::invalid_argument(f.method + ": " + why);
        pearl_fill rec{};
        rec.method = m->method_code;
        for (std::size_t i = 0; i < f.p.size(); ++i) rec.p[i] = f.p[i];
        auto direction = [&](int slot, double degrees) {
            const double a = degrees * 3.14159265358979323846 / 180.0;
            rec.dir[slot][0] = std::cos(a);
            rec.dir[slot][1] = std::sin(a);
        };
        if (rec.method == PEARL_FILL_YJUNCTION2D || rec.method == PEARL_FILL_YJUNCTIONLAMELLAE2D)
            for (int arm = 0; arm < 3; ++arm) direction(arm, f.p[6 + arm]);
        else if (rec.method == PEARL_FILL_ADHEREDPEARLITEARM2D || rec.method == PEARL_FILL_PEARLITEELLIPSE2D)
            direction(0, f.p[2]);
        out.push_back(rec);
    }
    return out;
}

}  // namespace pearl_host
