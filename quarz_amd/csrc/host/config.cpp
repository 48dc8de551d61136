#include "config.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <sstream>
#include <stdexcept>

namespace qz {

namespace {

std::string strip(const std::string& s) {
    size_t b = 0, e = s.size();
    while (b < e && std::isspace(static_cast<unsigned char>(s[b]))) ++b;
    while (e > b && std::isspace(static_cast<unsigned char>(s[e - 1]))) --e;
    return s.substr(b, e - b);
}

std::vector<std::string> words_of(const std::string& s) {
    std::istringstream in(s);
    std::vector<std::string> w;
    for (std::string t; in >> t;) w.push_back(t);
    return w;
}

// One pass over the deck text.  The scanner cuts it into events: a statement ("key = value"), the header of a
// block ("name {", "bunch name {", ...) and the end of a block ("}").  Statements end at a newline or at a ';';
// '#' starts a comment that runs to the end of the line; quotes (' or ") protect ';', '{' and '}'.  A '{' opens a
// block when no '=' precedes it in the statement; a '}' closes one when it is not balanced by a '{' of the same
// value.
struct Event {
    enum Kind { Statement, Open, Close } kind;
    std::string text;
    int line;
};

std::vector<Event> scan(const std::string& deck) {
    std::vector<Event> ev;
    std::string cur;
    int line = 1;
    char quote = 0;
    bool comment = false, has_eq = false;
    int depth = 0;   // braces opened inside the current value
    auto flush = [&]() {
        if (!strip(cur).empty()) ev.push_back({Event::Statement, cur, line});
        cur.clear();
        has_eq = false;
        depth = 0;
    };
    for (char c : deck) {
        if (c == '\n') {
            flush();
            comment = false;
            quote = 0;   // a quote does not continue past the end of its line
            ++line;
            continue;
        }
        if (comment) continue;
        if (quote) {
            if (c == quote) quote = 0;
            cur += c;
            continue;
        }
        switch (c) {
            case '#': comment = true; break;
            case '"':
            case '\'': quote = c; cur += c; break;
            case ';': flush(); break;
            case '=': has_eq = true; cur += c; break;
            case '{':
                if (has_eq) {
                    ++depth;
                    cur += c;
                } else {
                    ev.push_back({Event::Open, cur, line});
                    cur.clear();
                }
                break;
            case '}':
                if (has_eq && depth > 0) {
                    --depth;
                    cur += c;
                } else {
                    flush();
                    ev.push_back({Event::Close, std::string(), line});
                }
                break;
            default: cur += c;
        }
    }
    flush();
    return ev;
}

}  // namespace

void Config::parse(const std::string& text) {
    // prefixes of the open blocks, innermost last, and the list each named block joins
    std::vector<std::string> prefix;
    std::map<std::string, int> defined_at;
    std::vector<std::pair<std::string, std::string>> joins;   // (list key, name)
    for (const Event& e : scan(text)) {
        auto error = [&](const std::string& what) {
            throw std::runtime_error("input deck, line " + std::to_string(e.line) + ": " + what);
        };
        if (e.kind == Event::Close) {
            if (prefix.empty()) error("a '}' closes no block");
            prefix.pop_back();
            continue;
        }
        if (e.kind == Event::Open) {
            const std::vector<std::string> w = words_of(e.text);
            const std::string outer = prefix.empty() ? std::string() : prefix.back();
            // named blocks of the VLPL-family input: only at the top level
            static const std::map<std::string, std::pair<std::string, std::string>> named = {
                {"bunch", {"beams", ""}}, {"beam", {"beams", ""}}, {"species", {"plasma.species", ""}},
                {"diag", {"diag.names", "diag."}}};
            if (w.size() == 1) {
                prefix.push_back(outer + w[0] + ".");
            } else if (w.size() == 2 && outer.empty() && named.count(w[0])) {
                const auto& nb = named.at(w[0]);
                joins.emplace_back(nb.first, w[1]);
                prefix.push_back(nb.second + w[1] + ".");
            } else {
                error("cannot read the block header '" + strip(e.text) + "'");
            }
            continue;
        }
        const size_t eq = e.text.find('=');
        if (eq == std::string::npos) error("'" + strip(e.text) + "' is not of the form key = value");
        const std::string k = strip(e.text.substr(0, eq));
        if (k.empty()) error("a statement without a key");
        const std::string key = (prefix.empty() ? std::string() : prefix.back()) + k;
        const auto seen = defined_at.find(key);
        if (seen != defined_at.end())
            error("'" + key + "' is set a second time (it was set on line " + std::to_string(seen->second) + ")");
        defined_at[key] = e.line;
        kv_[key] = strip(e.text.substr(eq + 1));
    }
    if (!prefix.empty()) throw std::runtime_error("input deck: the block '" + prefix.back() + "' has no closing '}'");
    // a named block adds its name to its list, once
    for (const auto& j : joins) {
        std::string& list = kv_[j.first];
        const std::vector<std::string> have = words_of(list);
        if (std::find(have.begin(), have.end(), j.second) == have.end()) list = strip(list + " " + j.second);
    }
}

void Config::overrides(const std::string& text) {
    std::istringstream in(text);
    for (std::string a; std::getline(in, a);) {
        a = strip(a);
        if (a.empty() || a.compare(0, 2, "--") == 0) continue;   // option flags are not keys
        const size_t eq = a.find('=');
        if (eq == std::string::npos) continue;
        kv_[strip(a.substr(0, eq))] = strip(a.substr(eq + 1));
    }
}

bool Config::has(const std::string& key) const {
    return kv_.count(key) > 0 || (key == "units.n0_cm3" && kv_.count("units.n0") > 0);
}

namespace {
// leading number of `s` and the rest after it (the unit), or false when s does not start with a number
bool split_number(const std::string& s, double& x, std::string& unit) {
    const char* p = s.c_str();
    char* end = nullptr;
    x = std::strtod(p, &end);
    if (end == p) return false;
    unit = strip(std::string(end));
    return true;
}

// units a deck value may carry: dimension (L length, T time, N density, 1 plain factor) and the factor to SI
// (lengths m, times s, densities cm^-3) or to the plain number
struct UnitDef {
    char dim;
    double factor;
};
const std::map<std::string, UnitDef>& unit_table() {
    static const std::map<std::string, UnitDef> t = {
        // lengths
        {"nm", {'L', 1e-9}}, {"um", {'L', 1e-6}}, {"\xC2\xB5m", {'L', 1e-6}}, {"micron", {'L', 1e-6}},
        {"mm", {'L', 1e-3}}, {"cm", {'L', 1e-2}}, {"m", {'L', 1.0}}, {"mm*mrad", {'L', 1e-6}},
        {"mm.mrad", {'L', 1e-6}},
        // times
        {"as", {'T', 1e-18}}, {"fs", {'T', 1e-15}}, {"ps", {'T', 1e-12}}, {"ns", {'T', 1e-9}}, {"s", {'T', 1.0}},
        // densities
        {"cm^-3", {'N', 1.0}}, {"cm-3", {'N', 1.0}}, {"1/cm3", {'N', 1.0}}, {"1/cm^3", {'N', 1.0}},
        {"m^-3", {'N', 1e-6}}, {"m-3", {'N', 1e-6}}, {"1/m3", {'N', 1e-6}}, {"1/m^3", {'N', 1e-6}},
        // angles, percent and the normalized units themselves
        {"rad", {'1', 1.0}}, {"mrad", {'1', 1e-3}}, {"urad", {'1', 1e-6}}, {"deg", {'1', 0.017453292519943295}},
        {"%", {'1', 0.01}}, {"1/kp", {'1', 1.0}}, {"kp^-1", {'1', 1.0}}, {"c/wp", {'1', 1.0}}, {"1/wp", {'1', 1.0}},
        {"wp^-1", {'1', 1.0}}, {"n0", {'1', 1.0}}};
    return t;
}
}  // namespace

// The reference density: units.n0_cm3 (a plain number of cm^-3) or units.n0 (a density with its unit); 0 when
// neither is given.
double Config::n0_cm3() const {
    auto it = kv_.find("units.n0_cm3");
    if (it != kv_.end()) {
        used_.insert(it->first);
        double x;
        std::string u;
        if (!split_number(it->second, x, u) || !u.empty())
            throw std::runtime_error("units.n0_cm3 takes a plain number of particles per cm^3, not '" + it->second + "'");
        return x;
    }
    it = kv_.find("units.n0");
    if (it == kv_.end()) return 0.0;
    used_.insert(it->first);
    double x;
    std::string u;
    const auto& tab = unit_table();
    if (!split_number(it->second, x, u) || !tab.count(u) || tab.at(u).dim != 'N')
        throw std::runtime_error("units.n0 takes a density with its unit (cm^-3 or m^-3), not '" + it->second + "'");
    return x * tab.at(u).factor;
}

// A value in the code's normalized units: lengths in c/omega_p, times in 1/omega_p, densities in n0.  A unit
// converts with the plasma frequency of the reference density, omega_p = 5.64146e4 sqrt(n0 / cm^-3) rad/s.
double Config::number(const std::string& text, const std::string& what) const {
    double x;
    std::string u;
    if (!split_number(strip(text), x, u)) throw std::runtime_error(what + ": '" + text + "' is not a number");
    if (u.empty()) return x;
    const auto& tab = unit_table();
    const auto it = tab.find(u);
    if (it == tab.end()) throw std::runtime_error(what + ": '" + u + "' is not a unit this reader knows (in '" + text + "')");
    const UnitDef& d = it->second;
    if (d.dim == '1') return x * d.factor;
    const double n0 = n0_cm3();
    if (!(n0 > 0)) throw std::runtime_error(what + ": '" + text + "' has a unit, so units.n0 must give the reference density");
    if (d.dim == 'N') return x * d.factor / n0;
    const double omega_p = 5.64146e4 * std::sqrt(n0);   // rad/s
    if (d.dim == 'T') return x * d.factor * omega_p;
    return x * d.factor * omega_p / 2.99792458e8;       // lengths: k_p = omega_p / c
}

const std::string& Config::raw(const std::string& key) const {
    const auto it = kv_.find(key);
    if (it == kv_.end()) throw std::runtime_error("the input deck does not set '" + key + "', which this run needs");
    used_.insert(key);
    return it->second;
}

double Config::get_double(const std::string& key) const {
    if (key == "units.n0_cm3" && !kv_.count(key)) return n0_cm3();
    return number(raw(key), key);
}
double Config::get_double(const std::string& key, double def) const { return has(key) ? get_double(key) : def; }

int Config::get_int(const std::string& key) const {
    const std::string& v = raw(key);
    char* end = nullptr;
    const long x = std::strtol(v.c_str(), &end, 10);
    if (end == v.c_str() || !strip(end).empty()) throw std::runtime_error(key + ": '" + v + "' is not an integer");
    return static_cast<int>(x);
}
int Config::get_int(const std::string& key, int def) const { return has(key) ? get_int(key) : def; }

bool Config::get_bool(const std::string& key, bool def) const {
    if (!has(key)) return def;
    std::string v = raw(key);
    for (char& c : v) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    static const std::map<std::string, bool> words = {{"1", true},  {"true", true},   {"yes", true}, {"on", true},
                                                      {"0", false}, {"false", false}, {"no", false}, {"off", false}};
    const auto it = words.find(v);
    if (it == words.end()) throw std::runtime_error(key + ": '" + v + "' is neither true nor false");
    return it->second;
}

std::string Config::get_string(const std::string& key) const { return raw(key); }
std::string Config::get_string(const std::string& key, const std::string& def) const {
    return has(key) ? raw(key) : def;
}

std::vector<std::string> Config::get_list(const std::string& key) const {
    return has(key) ? words_of(raw(key)) : std::vector<std::string>();
}

std::vector<std::pair<double, double>> Config::get_pairs(const std::string& key) const {
    std::vector<std::pair<double, double>> out;
    for (const std::string& item : get_list(key)) {
        const size_t colon = item.find(':');
        if (colon == std::string::npos) throw std::runtime_error(key + ": '" + item + "' is not a pair a:b");
        out.emplace_back(number(item.substr(0, colon), key), number(item.substr(colon + 1), key));
    }
    return out;
}

std::string Config::dump() const {
    std::string out;
    for (const auto& kv : kv_) out += kv.first + " = " + kv.second + "\n";
    return out;
}

std::vector<std::string> Config::unused_keys() const {
    std::vector<std::string> out;
    for (const auto& kv : kv_)
        if (!used_.count(kv.first)) out.push_back(kv.first);
    return out;
}

}  // namespace qz
