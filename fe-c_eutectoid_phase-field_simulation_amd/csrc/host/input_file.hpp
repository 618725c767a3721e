// The input-file grammar of the reference program (read_config,
// pearl_serial.cpp:1600-1625): one `key = value` per line, `#` starts a
// comment, keys are unique, and keys starting with FILL carry a brace list
// of numbers and may repeat (one filling operation each, in file order).
// This reader only splits the file; the meaning of the keys lives in
// settings.hpp.  Texts of the errors are the reference's.
#pragma once

#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace pearl_host {

inline std::string strip(const std::string& s) {
    const char* ws = " \t\r\n";
    const std::size_t a = s.find_first_not_of(ws);
    return a == std::string::npos ? std::string() : s.substr(a, s.find_last_not_of(ws) - a + 1);
}

// A real number as std::stod reads it (the whole text, decimal or hex, no
// range error), or false.
inline bool parse_real(const std::string& text, double& out, std::size_t* used = nullptr) {
    const char* b = text.c_str();
    char* e = nullptr;
    errno = 0;
    const double v = std::strtod(b, &e);
    if (e == b || errno == ERANGE) return false;
    out = v;
    if (used) *used = static_cast<std::size_t>(e - b);
    return true;
}

// One FILL line: the method (the key) and its parameters.
struct FillLine {
    std::string method;
    std::vector<double> p;
};

class InputFile {
public:
    static InputFile read(const std::string& path) {
        std::ifstream in(path);
        if (!in) throw std::runtime_error("cannot open input: " + path);
        InputFile f;
        std::string raw;
        for (int number = 1; std::getline(in, raw); ++number) {
            const std::string line = strip(raw.substr(0, raw.find('#')));
            if (line.empty()) continue;
            const std::size_t eq = line.find('=');
            if (eq == std::string::npos) throw std::runtime_error("missing = on line " + std::to_string(number));
            const std::string key = strip(line.substr(0, eq)), value = strip(line.substr(eq + 1));
            if (key.compare(0, 4, "FILL") == 0) {
                f.fills_.push_back({key, brace_list(key, value, number)});
            } else if (key.empty() || value.empty() || !f.values_.emplace(key, Value{value, false}).second) {
                throw std::runtime_error("empty/duplicate key: " + key);
            }
        }
        return f;
    }

    bool has(const std::string& key) const { return values_.count(key) != 0; }

    // The text of `key`, marked as used; nullptr when absent.
    const std::string* take(const std::string& key) {
        auto it = values_.find(key);
        if (it == values_.end()) return nullptr;
        it->second.used = true;
        return &it->second.text;
    }

    const std::vector<FillLine>& fills() const { return fills_; }

    // Every key must have been read by someone (the first unread key, in
    // key order, is reported).
    void reject_unread() const {
        for (const auto& kv : values_)
            if (!kv.second.used) throw std::runtime_error("unknown key: " + kv.first);
    }

private:
    struct Value {
        std::string text;
        bool used;
    };
    std::map<std::string, Value> values_;
    std::vector<FillLine> fills_;

    // "{v1, v2, ...}" with an optional trailing ';', every entry a finite number.
    static std::vector<double> brace_list(const std::string& method, const std::string& value, int line) {
        const std::invalid_argument bad("Invalid " + method + " filling parameters at input line " +
                                        std::to_string(line) + ": expected {finite_number,...}, optionally followed by ';'");
        std::string body = value;
        if (!body.empty() && body.back() == ';') body = strip(body.substr(0, body.size() - 1));
        if (body.size() < 3 || body.front() != '{' || body.back() != '}') throw bad;
        body = body.substr(1, body.size() - 2);
        std::vector<double> out;
        std::size_t from = 0;
        for (;;) {
            const std::size_t comma = body.find(',', from);
            const std::string item =
                strip(body.substr(from, comma == std::string::npos ? std::string::npos : comma - from));
            double v = 0;
            std::size_t used = 0;
            if (item.empty() || !parse_real(item, v, &used) || used != item.size() || !std::isfinite(v)) throw bad;
            out.push_back(v);
            if (comma == std::string::npos) return out;
            from = comma + 1;
        }
    }
};

}  // namespace pearl_host
