#include "config.hpp"

#include <algorithm>
#include <cmath>
#include <sstream>
#include <stdexcept>

namespace qz {

namespace {
std::string trim(const std::string& s) {
    const char* ws = " \t\r\n";
    const auto b = s.find_first_not_of(ws);
    if (b == std::string::npos) return "";
    const auto e = s.find_last_not_of(ws);
    return s.substr(b, e - b + 1);
}

// split at ';' outside quotes (src/Config.cpp:21-33)
std::vector<std::string> split_statements(const std::string& line) {
    std::vector<std::string> out(1);
    char quote = 0;
    for (char c : line) {
        if (quote) {
            if (c == quote) quote = 0;
            out.back() += c;
        } else if (c == '"' || c == '\'') {
            quote = c;
            out.back() += c;
        } else if (c == ';') {
            out.emplace_back();
        } else {
            out.back() += c;
        }
    }
    return out;
}

// '}' minus '{' outside quotes (src/Config.cpp:36-46)
int excess_closing(const std::string& s) {
    int n = 0;
    char quote = 0;
    for (char c : s) {
        if (quote) {
            if (c == quote) quote = 0;
        } else if (c == '"' || c == '\'') {
            quote = c;
        } else if (c == '}') {
            ++n;
        } else if (c == '{') {
            --n;
        }
    }
    return n;
}
}  // namespace

void Config::parse(const std::string& text) {
    std::istringstream in(text);
    std::string line;
    int lineno = 0;
    std::vector<std::string> blocks;   // key prefixes of the open blocks
    std::map<std::string, int> first_line;
    std::vector<std::pair<std::string, std::string>> lists;   // named blocks -> their list key
    auto fail = [&](const std::string& m) {
        throw std::runtime_error("Config: line " + std::to_string(lineno) + ": " + m);
    };
    auto pop = [&]() {
        if (blocks.empty()) fail("'}' without an open block");
        blocks.pop_back();
    };
    while (std::getline(in, line)) {
        ++lineno;
        const auto hash = line.find('#');
        if (hash != std::string::npos) line.erase(hash);
        for (std::string st : split_statements(line)) {
            st = trim(st);
            while (!st.empty()) {
                if (st[0] == '}') {
                    pop();
                    st = trim(st.substr(1));
                    continue;
                }
                const auto eq = st.find('=');
                const auto brace = st.find('{');
                if (brace != std::string::npos && (eq == std::string::npos || brace < eq)) {
                    std::istringstream hs(st.substr(0, brace));
                    std::vector<std::string> words;
                    for (std::string t; hs >> t;) words.push_back(t);
                    const std::string outer = blocks.empty() ? std::string() : blocks.back();
                    if (words.size() == 1) {
                        blocks.push_back(outer + words[0] + ".");
                    } else if (words.size() == 2 && outer.empty() && (words[0] == "bunch" || words[0] == "beam")) {
                        lists.emplace_back("beams", words[1]);
                        blocks.push_back(words[1] + ".");
                    } else if (words.size() == 2 && outer.empty() && words[0] == "species") {
                        lists.emplace_back("plasma.species", words[1]);
                        blocks.push_back(words[1] + ".");
                    } else if (words.size() == 2 && outer.empty() && words[0] == "diag") {
                        lists.emplace_back("diag.names", words[1]);
                        blocks.push_back("diag." + words[1] + ".");
                    } else {
                        fail("block header '" + trim(st.substr(0, brace)) + " {' not understood");
                    }
                    st = trim(st.substr(brace + 1));
                    continue;
                }
                if (eq == std::string::npos) fail("missing '=' in '" + st + "'");
                std::string key = trim(st.substr(0, eq));
                std::string val = trim(st.substr(eq + 1));
                if (key.empty()) fail("empty key");
                int closes = 0;
                while (!blocks.empty() && !val.empty() && val.back() == '}' && excess_closing(val) > 0) {
                    val = trim(val.substr(0, val.size() - 1));
                    ++closes;
                }
                if (!blocks.empty()) key = blocks.back() + key;
                if (first_line.count(key))
                    fail("key '" + key + "' is given twice (first in line " + std::to_string(first_line[key]) + ")");
                first_line[key] = lineno;
                kv_[key] = val;
                for (int i = 0; i < closes; ++i) pop();
                st.clear();
            }
        }
    }
    if (!blocks.empty()) throw std::runtime_error("Config: a block is not closed");
    for (const auto& l : lists) {
        std::istringstream ls(kv_.count(l.first) ? kv_[l.first] : std::string());
        bool present = false;
        for (std::string t; ls >> t;) present = present || t == l.second;
        if (!present) kv_[l.first] += (kv_[l.first].empty() ? "" : " ") + l.second;
    }
}

void Config::overrides(const std::string& text) {
    std::istringstream in(text);
    std::string a;
    while (std::getline(in, a)) {
        a = trim(a);
        const auto eq = a.find('=');
        if (a.empty() || eq == std::string::npos || a.rfind("--", 0) == 0) continue;
        kv_[trim(a.substr(0, eq))] = trim(a.substr(eq + 1));
    }
}

bool Config::has(const std::string& key) const {
    if (kv_.count(key)) return true;
    return key == "units.n0_cm3" && kv_.count("units.n0");
}

double Config::n0_cm3() const {
    const bool plain = kv_.count("units.n0_cm3") > 0;
    if (!plain && !kv_.count("units.n0")) return 0.0;
    const std::string& v = raw(plain ? "units.n0_cm3" : "units.n0");
    size_t pos = 0;
    double x = 0;
    try { x = std::stod(v, &pos); } catch (...) { pos = 0; }
    const std::string u = pos ? trim(v.substr(pos)) : std::string("?");
    if (plain) {
        if (!u.empty()) throw std::runtime_error("Config: key 'units.n0_cm3' is a plain number of cm^-3: " + v);
        return x;
    }
    if (u == "cm^-3" || u == "cm-3" || u == "1/cm3" || u == "1/cm^3") return x;
    if (u == "m^-3" || u == "m-3" || u == "1/m3" || u == "1/m^3") return x * 1e-6;
    throw std::runtime_error("Config: key 'units.n0' needs a density with its unit (cm^-3 or m^-3): " + v);
}

// numbers with units (src/Config.cpp:149-183): lengths in 1/k_p, times in 1/omega_p, densities in n0
double Config::number(const std::string& text, const std::string& what) const {
    const std::string v = trim(text);
    size_t pos = 0;
    double x = 0;
    try { x = std::stod(v, &pos); } catch (...) { pos = 0; }
    if (pos == 0) throw std::runtime_error("Config: " + what + " is not a number: " + text);
    const std::string u = trim(v.substr(pos));
    if (u.empty()) return x;
    struct Unit { const char* name; char kind; double si; };
    static const Unit units[] = {
        {"nm", 'L', 1e-9}, {"um", 'L', 1e-6}, {"\xC2\xB5m", 'L', 1e-6}, {"micron", 'L', 1e-6}, {"mm", 'L', 1e-3},
        {"cm", 'L', 1e-2}, {"m", 'L', 1.0}, {"mm*mrad", 'L', 1e-6}, {"mm.mrad", 'L', 1e-6},
        {"as", 'T', 1e-18}, {"fs", 'T', 1e-15}, {"ps", 'T', 1e-12}, {"ns", 'T', 1e-9}, {"s", 'T', 1.0},
        {"cm^-3", 'N', 1.0}, {"cm-3", 'N', 1.0}, {"1/cm3", 'N', 1.0}, {"1/cm^3", 'N', 1.0},
        {"m^-3", 'N', 1e-6}, {"m-3", 'N', 1e-6}, {"1/m3", 'N', 1e-6}, {"1/m^3", 'N', 1e-6},
        {"rad", '1', 1.0}, {"mrad", '1', 1e-3}, {"urad", '1', 1e-6}, {"deg", '1', 0.017453292519943295},
        {"%", '1', 0.01},
        {"1/kp", '1', 1.0}, {"kp^-1", '1', 1.0}, {"c/wp", '1', 1.0}, {"1/wp", '1', 1.0}, {"wp^-1", '1', 1.0},
        {"n0", '1', 1.0},
    };
    for (const Unit& un : units) {
        if (u != un.name) continue;
        if (un.kind == '1') return x * un.si;
        const double n0 = n0_cm3();
        if (!(n0 > 0))
            throw std::runtime_error("Config: " + what + " = " + text + ": a value with a unit needs units.n0");
        if (un.kind == 'N') return x * un.si / n0;
        const double wp = 5.64146e4 * std::sqrt(n0);
        if (un.kind == 'T') return x * un.si * wp;
        return x * un.si * wp / 2.99792458e8;
    }
    throw std::runtime_error("Config: " + what + " = " + text + ": unknown unit '" + u + "'");
}

const std::string& Config::raw(const std::string& key) const {
    const auto it = kv_.find(key);
    if (it == kv_.end()) throw std::runtime_error("Config: required key '" + key + "' is missing");
    used_.insert(key);
    return it->second;
}

double Config::get_double(const std::string& key) const {
    if (key == "units.n0_cm3" && !kv_.count(key) && kv_.count("units.n0")) return n0_cm3();
    return number(raw(key), "key '" + key + "'");
}
double Config::get_double(const std::string& key, double def) const { return has(key) ? get_double(key) : def; }
int Config::get_int(const std::string& key) const {
    const std::string& v = raw(key);
    try {
        return static_cast<int>(std::stol(v));
    } catch (...) {
        throw std::runtime_error("Config: key '" + key + "' is not an integer: " + v);
    }
}
int Config::get_int(const std::string& key, int def) const { return has(key) ? get_int(key) : def; }
bool Config::get_bool(const std::string& key, bool def) const {
    if (!has(key)) return def;
    std::string v = raw(key);
    std::transform(v.begin(), v.end(), v.begin(), ::tolower);
    if (v == "1" || v == "true" || v == "yes" || v == "on") return true;
    if (v == "0" || v == "false" || v == "no" || v == "off") return false;
    throw std::runtime_error("Config: key '" + key + "' is not a boolean: " + v);
}
std::string Config::get_string(const std::string& key) const { return raw(key); }
std::string Config::get_string(const std::string& key, const std::string& def) const {
    return has(key) ? raw(key) : def;
}
std::vector<std::string> Config::get_list(const std::string& key) const {
    std::vector<std::string> out;
    if (!has(key)) return out;
    std::istringstream ss(raw(key));
    for (std::string tok; ss >> tok;) out.push_back(tok);
    return out;
}
std::vector<std::pair<double, double>> Config::get_pairs(const std::string& key) const {
    std::vector<std::pair<double, double>> out;
    for (const auto& tok : get_list(key)) {
        const auto c = tok.find(':');
        if (c == std::string::npos) throw std::runtime_error("Config: key '" + key + "': expected a:b pairs, got '" + tok + "'");
        out.emplace_back(number(tok.substr(0, c), "key '" + key + "'"), number(tok.substr(c + 1), "key '" + key + "'"));
    }
    return out;
}
std::vector<std::string> Config::unused_keys() const {
    std::vector<std::string> out;
    for (const auto& kv : kv_)
        if (!used_.count(kv.first)) out.push_back(kv.first);
    return out;
}

}  // namespace qz
