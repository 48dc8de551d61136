// ORACLE -- test infrastructure, not product code.
//
// A CPU restatement of the QUARZ quasi-static time step (per-slice plasma sweep, beam deposit and
// Vay / Higuera-Cary beam push), written from the reference's source as text and following its
// arithmetic and expression order literally, so that a correct gfx950 implementation can be compared
// with it bit for bit.  Only the tests, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
// load this library; the product (quarz_amd) never links, calls or imports it.
//
// Each function cites the reference file:line it restates (paths relative to the reference checkout).
//
// Two execution modes share one source:
//   * serial  (PAR = false): one thread, deposits added in particle-index order.  This is the order of
//     the reference's Kokkos::Serial backend and the order the GPU path reproduces exactly.  It is the
//     checker.
//   * OpenMP  (PAR = true):  the reference's loop structure (src/Simulation.cpp:394-819,
//     src/PlasmaSpecies.cpp:409-668): one parallel loop per kernel, `omp atomic` deposits, serial
//     Thomas solves.  Atomics make the summation order (and the last bits) run-dependent, exactly as in
//     the reference's OpenMP build (README.md:1498).  It is timed as the CPU baseline; it is never used
//     as a checker.
//
// Compiled with g++ -O3 -ffp-contract=off (no -march=native): a default x86-64 build without FMA, as a
// default Release build of the reference.
//
// Parity pins (tests/test_oracle_*.py): Thomas on random diagonally dominant systems
// (tests/test_solvers.cpp:34-56), manufactured-solution orders of L0/L1F/L1D/L2D (:58-124), the uniform
// deposit (:126-143), the AB coefficients (tests/test_ab.cpp:17-43), the pusher identities and the printed
// numbers of validation/pusher_test_output.txt, the std::mt19937_64 known answer, hat volumes in closed
// form, the blowout / linear-wake validation numbers of validation/reference_full.txt:1-8, the m = 1 shift
// validation with a rigid analytic driver (reference_full.txt:20-24), the 90-degree rotation covariance of the
// m = 1 equations, and the mobile-ion validation of reference_quick.txt:12-16.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <random>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

using Real = double;
constexpr Real PI = 3.14159265358979323846;   // src/Types.hpp:22

// glibc pow/exp called through a pointer, so that neither compiler folds a call with constant
// arguments at build time: the oracle and the product both take these values from the C library at run
// time and therefore agree bit for bit.
static double (*volatile libm_pow)(double, double) = ::pow;

// xi user <-> internal (src/Types.hpp:30-31)
inline Real xi_from_user(Real xu) { return Real(0) - xu; }
inline Real xi_to_user(Real xi) { return Real(0) - xi; }

static std::string last_error;

// ============================================================================ Config (src/Config.cpp)
static std::string trim(const std::string& s) {
    const char* ws = " \t\r\n";
    auto b = s.find_first_not_of(ws);
    if (b == std::string::npos) return "";
    auto e = s.find_last_not_of(ws);
    return s.substr(b, e - b + 1);
}

// src/Config.cpp:21-33
static std::vector<std::string> statements(const std::string& line) {
    std::vector<std::string> out;
    std::string cur;
    char quote = 0;
    for (char c : line) {
        if (quote) { if (c == quote) quote = 0; cur += c; }
        else if (c == '"' || c == '\'') { quote = c; cur += c; }
        else if (c == ';') { out.push_back(cur); cur.clear(); }
        else cur += c;
    }
    out.push_back(cur);
    return out;
}

// src/Config.cpp:36-46
static int closing_excess(const std::string& s) {
    int n = 0;
    char quote = 0;
    for (char c : s) {
        if (quote) { if (c == quote) quote = 0; }
        else if (c == '"' || c == '\'') quote = c;
        else if (c == '}') ++n;
        else if (c == '{') --n;
    }
    return n;
}

struct Config {
    std::map<std::string, std::string> kv;

    // src/Config.cpp:48-125 (deck text instead of a file name)
    void parse(const std::string& text) {
        std::istringstream in(text);
        std::string line;
        int lineno = 0;
        std::vector<std::string> prefix;
        std::map<std::string, int> where;
        std::vector<std::pair<std::string, std::string>> named;
        auto err = [&](const std::string& m) {
            throw std::runtime_error("Config: line " + std::to_string(lineno) + ": " + m);
        };
        auto close = [&]() {
            if (prefix.empty()) err("'}' without an open block");
            prefix.pop_back();
        };
        while (std::getline(in, line)) {
            ++lineno;
            auto hash = line.find('#');
            if (hash != std::string::npos) line = line.substr(0, hash);
            for (std::string s : statements(line)) {
                s = trim(s);
                while (!s.empty()) {
                    if (s[0] == '}') { close(); s = trim(s.substr(1)); continue; }
                    const auto eq = s.find('=');
                    const auto br = s.find('{');
                    if (br != std::string::npos && (eq == std::string::npos || br < eq)) {
                        std::istringstream hs(s.substr(0, br));
                        std::vector<std::string> w;
                        for (std::string t; hs >> t;) w.push_back(t);
                        const std::string outer = prefix.empty() ? std::string() : prefix.back();
                        if (w.size() == 1) prefix.push_back(outer + w[0] + ".");
                        else if (w.size() == 2 && outer.empty() && (w[0] == "bunch" || w[0] == "beam")) {
                            named.emplace_back("beams", w[1]);
                            prefix.push_back(w[1] + ".");
                        } else if (w.size() == 2 && outer.empty() && w[0] == "species") {
                            named.emplace_back("plasma.species", w[1]);
                            prefix.push_back(w[1] + ".");
                        } else if (w.size() == 2 && outer.empty() && w[0] == "diag") {
                            named.emplace_back("diag.names", w[1]);
                            prefix.push_back("diag." + w[1] + ".");
                        } else
                            err("bad block header");
                        s = trim(s.substr(br + 1));
                        continue;
                    }
                    if (eq == std::string::npos) err("missing '=' in '" + s + "'");
                    std::string key = trim(s.substr(0, eq));
                    std::string val = trim(s.substr(eq + 1));
                    if (key.empty()) err("empty key");
                    int nclose = 0;
                    while (!prefix.empty() && !val.empty() && val.back() == '}' && closing_excess(val) > 0) {
                        val = trim(val.substr(0, val.size() - 1));
                        ++nclose;
                    }
                    if (!prefix.empty()) key = prefix.back() + key;
                    if (where.count(key)) err("key '" + key + "' is given twice");
                    where[key] = lineno;
                    kv[key] = val;
                    for (int i = 0; i < nclose; ++i) close();
                    s.clear();
                }
            }
        }
        if (!prefix.empty()) throw std::runtime_error("Config: a block is not closed");
        for (const auto& nl : named) {
            std::istringstream ls(kv.count(nl.first) ? kv[nl.first] : std::string());
            bool have = false;
            for (std::string t; ls >> t;) have = have || t == nl.second;
            if (!have) kv[nl.first] += (kv[nl.first].empty() ? "" : " ") + nl.second;
        }
    }
    // command-line style overrides "k=v" separated by newlines (src/main.cpp:82-93)
    void overrides(const std::string& text) {
        std::istringstream in(text);
        std::string a;
        while (std::getline(in, a)) {
            a = trim(a);
            auto eq = a.find('=');
            if (a.empty() || eq == std::string::npos) continue;
            kv[trim(a.substr(0, eq))] = trim(a.substr(eq + 1));
        }
    }
    // src/Config.cpp:127-130
    bool has(const std::string& key) const {
        if (kv.count(key)) return true;
        return key == "units.n0_cm3" && kv.count("units.n0");
    }
    // src/Config.cpp:132-147
    double n0_cm3() const {
        const bool plain = kv.count("units.n0_cm3") > 0;
        if (!plain && !kv.count("units.n0")) return 0.0;
        const std::string& v = kv.at(plain ? "units.n0_cm3" : "units.n0");
        size_t pos = 0;
        double x = 0;
        try { x = std::stod(v, &pos); } catch (...) { pos = 0; }
        const std::string u = pos ? trim(v.substr(pos)) : std::string("?");
        if (plain) {
            if (!u.empty()) throw std::runtime_error("Config: units.n0_cm3 is a plain number of cm^-3");
            return x;
        }
        if (u == "cm^-3" || u == "cm-3" || u == "1/cm3" || u == "1/cm^3") return x;
        if (u == "m^-3" || u == "m-3" || u == "1/m3" || u == "1/m^3") return x * 1e-6;
        throw std::runtime_error("Config: units.n0 needs a density with its unit");
    }
    // src/Config.cpp:149-183
    double number(const std::string& text, const std::string& what) const {
        const std::string v = trim(text);
        size_t pos = 0;
        double x = 0;
        try { x = std::stod(v, &pos); } catch (...) { pos = 0; }
        if (pos == 0) throw std::runtime_error("Config: " + what + " is not a number: " + text);
        const std::string u = trim(v.substr(pos));
        if (u.empty()) return x;
        struct U { const char* name; char kind; double si; };
        static const U table[] = {
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
        for (const U& t : table) {
            if (u != t.name) continue;
            if (t.kind == '1') return x * t.si;
            const double n0 = n0_cm3();
            if (!(n0 > 0)) throw std::runtime_error("Config: " + what + ": a value with a unit needs units.n0");
            if (t.kind == 'N') return x * t.si / n0;
            const double wp = 5.64146e4 * std::sqrt(n0);
            if (t.kind == 'T') return x * t.si * wp;
            return x * t.si * wp / 2.99792458e8;
        }
        throw std::runtime_error("Config: " + what + " = " + text + ": unknown unit '" + u + "'");
    }
    const std::string& raw(const std::string& key) const {
        auto it = kv.find(key);
        if (it == kv.end()) throw std::runtime_error("Config: required key '" + key + "' is missing");
        return it->second;
    }
    double get_double(const std::string& key) const {
        if (key == "units.n0_cm3" && !kv.count(key) && kv.count("units.n0")) return n0_cm3();
        return number(raw(key), "key '" + key + "'");
    }
    double get_double(const std::string& key, double def) const { return has(key) ? get_double(key) : def; }
    int get_int(const std::string& key) const {
        const std::string& v = raw(key);
        try { return static_cast<int>(std::stol(v)); }
        catch (...) { throw std::runtime_error("Config: key '" + key + "' is not an integer: " + v); }
    }
    int get_int(const std::string& key, int def) const { return has(key) ? get_int(key) : def; }
    bool get_bool(const std::string& key, bool def) const {
        if (!has(key)) return def;
        std::string v = raw(key);
        std::transform(v.begin(), v.end(), v.begin(), ::tolower);
        if (v == "1" || v == "true" || v == "yes" || v == "on") return true;
        if (v == "0" || v == "false" || v == "no" || v == "off") return false;
        throw std::runtime_error("Config: key '" + key + "' is not a boolean: " + v);
    }
    std::string get_string(const std::string& key) const { return raw(key); }
    std::string get_string(const std::string& key, const std::string& def) const { return has(key) ? raw(key) : def; }
    std::vector<std::string> get_list(const std::string& key) const {
        std::vector<std::string> out;
        if (!has(key)) return out;
        std::istringstream ss(raw(key));
        std::string tok;
        while (ss >> tok) out.push_back(tok);
        return out;
    }
    std::vector<std::pair<double, double>> get_pairs(const std::string& key) const {
        std::vector<std::pair<double, double>> out;
        for (const auto& tok : get_list(key)) {
            auto c = tok.find(':');
            if (c == std::string::npos) throw std::runtime_error("Config: key '" + key + "': expected a:b pairs");
            out.emplace_back(number(tok.substr(0, c), key), number(tok.substr(c + 1), key));
        }
        return out;
    }
};

// ============================================================================ radial grid (src/RadialGrid.*)
// src/RadialGrid.cpp:14-20
static std::vector<Real> make_uniform(Real R, Real dr) {
    if (R <= 0 || dr <= 0) throw std::runtime_error("grid: R and dr must be positive");
    const int n = std::max(1, static_cast<int>(std::lround(R / dr)));
    std::vector<Real> nodes(n + 1);
    for (int j = 0; j <= n; ++j) nodes[j] = R * Real(j) / Real(n);
    return nodes;
}

// src/RadialGrid.cpp:27-71
static std::vector<Real> make_regions(Real R, const std::vector<std::pair<double, double>>& regions, Real max_ratio) {
    if (regions.empty()) throw std::runtime_error("grid: grid.regions is empty");
    if (max_ratio < 1.0) throw std::runtime_error("grid: grid.max_ratio must be >= 1");
    for (size_t i = 0; i < regions.size(); ++i) {
        if (regions[i].second <= 0) throw std::runtime_error("grid: region spacing must be > 0");
        if (i > 0 && regions[i].first <= regions[i - 1].first)
            throw std::runtime_error("grid: region end radii must increase");
    }
    const Real g = max_ratio - 1.0;
    auto target = [&](Real x) {
        Real ht = regions.back().second;
        for (const auto& reg : regions)
            if (x < reg.first) { ht = reg.second; break; }
        Real start = 0;
        for (const auto& reg : regions) {
            if (start > x && g > 0) ht = std::min(ht, Real(reg.second + g * (start - x)));
            if (start > x && g == 0) ht = std::min(ht, Real(reg.second));
            start = reg.first;
        }
        return ht;
    };
    std::vector<Real> nodes{0.0};
    Real x = 0, hprev = target(0);
    while (x < R * (1 - 1e-12)) {
        Real ht = target(x + 0.5 * hprev);
        Real hnew = std::clamp(ht, hprev / max_ratio, hprev * max_ratio);
        if (nodes.size() == 1) hnew = target(0);
        x += hnew;
        nodes.push_back(x);
        hprev = hnew;
        if (nodes.size() > 10000000) throw std::runtime_error("grid: too many nodes");
    }
    const size_t n = nodes.size();
    if (n >= 3 && (nodes[n - 1] - R) > (R - nodes[n - 2])) nodes.pop_back();
    const Real scale = R / nodes.back();
    for (auto& v : nodes) v *= scale;
    nodes.back() = R;
    return nodes;
}

// src/RadialGrid.cpp:73-85 (one node per line, '#' comments)
static std::vector<Real> read_nodes_text(const std::string& text) {
    std::istringstream in(text);
    std::vector<Real> nodes;
    std::string line;
    while (std::getline(in, line)) {
        auto hash = line.find('#');
        if (hash != std::string::npos) line = line.substr(0, hash);
        if (line.find_first_not_of(" \t\r\n") == std::string::npos) continue;
        nodes.push_back(std::stod(line));
    }
    return nodes;
}

struct Grid {
    int N = 0;
    Real R = 0;
    std::vector<Real> r, h, rmid, V, inv_h, inv_V;

    // src/RadialGrid.cpp:106-148
    void finalize(const std::vector<Real>& nodes) {
        if (nodes.size() < 3) throw std::runtime_error("grid: need at least 2 cells");
        if (std::abs(nodes[0]) > 0) throw std::runtime_error("grid: first node must be r=0");
        for (size_t j = 1; j < nodes.size(); ++j)
            if (!(nodes[j] > nodes[j - 1])) throw std::runtime_error("grid: nodes must be strictly increasing");
        r = nodes;
        N = static_cast<int>(r.size()) - 1;
        R = r[N];
        h.resize(N);
        rmid.resize(N);
        for (int j = 0; j < N; ++j) {
            h[j] = r[j + 1] - r[j];
            rmid[j] = 0.5 * (r[j + 1] + r[j]);
        }
        V.assign(N + 1, 0.0);
        for (int j = 0; j <= N; ++j) {
            Real v = 0;
            if (j > 0) v += r[j - 1] * h[j - 1] / 2 + h[j - 1] * h[j - 1] / 3;
            if (j < N) v += r[j] * h[j] / 2 + h[j] * h[j] / 6;
            V[j] = 2 * PI * v;
        }
        inv_h.resize(N);
        inv_V.resize(N + 1);
        for (int j = 0; j <= N; ++j) inv_V[j] = 1.0 / V[j];
        for (int j = 0; j < N; ++j) inv_h[j] = 1.0 / h[j];
    }
    // GridD::locate (src/RadialGrid.hpp:35-44)
    int locate(Real x) const {
        if (x <= Real(0)) return 0;
        if (x >= R) return N - 1;
        int lo = 0, hi = N;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (r[mid] <= x) lo = mid; else hi = mid;
        }
        return lo;
    }
    // GridD::weights (src/RadialGrid.hpp:46-50)
    void weights(Real x, int& j, Real& t) const {
        j = locate(x);
        t = (x - r[j]) * inv_h[j];
        t = t < Real(0) ? Real(0) : (t > Real(1) ? Real(1) : t);
    }
};

// RadialGrid(const Config&) (src/RadialGrid.cpp:88-104); grid.type = file reads grid.file_text here
static std::vector<Real> grid_nodes(const Config& cfg) {
    const std::string type = cfg.get_string("grid.type", "uniform");
    if (type == "uniform") return make_uniform(cfg.get_double("grid.rmax"), cfg.get_double("grid.dr"));
    if (type == "regions" || type == "stretched")
        return make_regions(cfg.get_double("grid.rmax"), cfg.get_pairs("grid.regions"), cfg.get_double("grid.max_ratio", 1.05));
    if (type == "file") {
        std::ifstream in(cfg.get_string("grid.file"));
        if (!in) throw std::runtime_error("grid: cannot open grid file");
        std::stringstream ss;
        ss << in.rdbuf();
        return read_nodes_text(ss.str());
    }
    throw std::runtime_error("grid: unknown grid.type '" + type + "'");
}

// ============================================================================ Thomas (src/Tridiag.cpp:111-126)
static void thomas(const Real* a, const Real* b, const Real* c, const Real* d, Real* x, int M, Real* cp, Real* dp) {
    cp[0] = c[0] / b[0];
    dp[0] = d[0] / b[0];
    for (int i = 1; i < M; ++i) {
        const Real m = Real(1) / (b[i] - a[i] * cp[i - 1]);
        cp[i] = c[i] * m;
        dp[i] = (d[i] - a[i] * dp[i - 1]) * m;
    }
    x[M - 1] = dp[M - 1];
    for (int i = M - 2; i >= 0; --i) x[i] = dp[i] - cp[i] * x[i + 1];
}

// ============================================================================ field solver (src/FieldSolver.cpp)
enum OpKind { L0 = 0, L1D = 1, L1F = 2, L2D = 3 };

struct Op { std::vector<Real> a, b, c; bool flux = false; bool axis_regular = false; };

// src/FieldSolver.cpp:20-33
static void fv_rows(const Grid& g, int n, std::vector<Real>& a, std::vector<Real>& b, std::vector<Real>& c) {
    const int N = g.N, M = N + 1;
    a.assign(M, 0); b.assign(M, 0); c.assign(M, 0);
    b[0] = 1.0;
    b[N] = 1.0;
    auto kp = [&](int j) { return 1.0 / (g.h[j] * libm_pow(g.rmid[j], 2 * n - 1)); };
    for (int j = 1; j < N; ++j) {
        const Real dV = 0.5 * (g.h[j] + g.h[j - 1]);
        const Real pre = libm_pow(g.r[j], n - 1) / dV;
        c[j] = pre * libm_pow(g.r[j + 1], n) * kp(j);
        a[j] = pre * libm_pow(g.r[j - 1], n) * kp(j - 1);
        b[j] = -pre * libm_pow(g.r[j], n) * (kp(j) + kp(j - 1));
    }
}

struct Solver {
    const Grid* g = nullptr;
    Op ops[4];
    Real flux_coef = 0;
    std::vector<Real> gm, g0, gp, ir;
    mutable std::vector<Real> bw, dw, cp, dp;

    // src/FieldSolver.cpp:43-109
    void init(const Grid& grid) {
        g = &grid;
        const int N = grid.N, M = N + 1;
        {
            std::vector<Real> a(M, 0), b(M, 0), c(M, 0);
            for (int j = 0; j < N; ++j) {
                const Real invV = 2 * PI / grid.V[j];
                a[j] = (j > 0) ? grid.rmid[j - 1] / grid.h[j - 1] * invV : 0.0;
                c[j] = grid.rmid[j] / grid.h[j] * invV;
                b[j] = -(a[j] + c[j]);
            }
            b[N] = 1.0;
            ops[L0].a = a; ops[L0].b = b; ops[L0].c = c;
            ops[L0].axis_regular = true;
        }
        auto build_fv = [&](int n, bool flux, Op& op) {
            std::vector<Real> a, b, c;
            fv_rows(grid, n, a, b, c);
            auto kp = [&](int j) { return 1.0 / (grid.h[j] * libm_pow(grid.rmid[j], 2 * n - 1)); };
            if (flux) {
                const Real dV = 0.5 * grid.h[N - 1];
                const Real k = kp(N - 1) / dV;
                a[N] = grid.r[N - 1] * k;
                b[N] = -grid.r[N] * k;
                flux_coef = 1.0 / dV;
            } else {
                b[N] = 1.0;
            }
            op.a = a; op.b = b; op.c = c;
            op.flux = flux;
            op.axis_regular = false;
        };
        build_fv(1, false, ops[L1D]);
        build_fv(1, true, ops[L1F]);
        build_fv(2, false, ops[L2D]);
        gm.assign(M, 0); g0.assign(M, 0); gp.assign(M, 0); ir.assign(M, 0);
        for (int j = 1; j < N; ++j) {
            const Real hm = grid.h[j - 1], hp = grid.h[j];
            const Real den = hm * hp * (hm + hp);
            gp[j] = hm * hm / den;
            gm[j] = -hp * hp / den;
            g0[j] = (hp * hp - hm * hm) / den;
        }
        {
            const Real h1 = grid.h[N - 1], h2 = (N >= 2) ? grid.h[N - 2] : grid.h[N - 1];
            const Real s = h1 + h2;
            g0[N] = (2 * h1 + h2) / (h1 * s);
            gm[N] = -s / (h1 * h2);
            gp[N] = h1 / (h2 * s);
        }
        for (int j = 1; j <= N; ++j) ir[j] = 1.0 / grid.r[j];
        ir[0] = 1.0 / grid.r[1];
        bw.assign(M, 0); dw.assign(M, 0); cp.assign(M, 0); dp.assign(M, 0);
    }

    // src/FieldSolver.cpp:242-261; chi/wall may be null
    void solve(int k, const Real* rhs, Real* f, const Real* chi = nullptr, const Real* wall = nullptr,
               Real scale = Real(1)) const {
        const Op& o = ops[k];
        const int N = g->N;
        for (int j = 0; j <= N; ++j) {
            const bool dirichlet_row = (j == N && !o.flux) || (j == 0 && !o.axis_regular);
            if (dirichlet_row) { bw[j] = o.b[j]; dw[j] = Real(0); continue; }
            bw[j] = chi ? o.b[j] - chi[j] : o.b[j];
            dw[j] = scale * rhs[j];
            if (j == N && o.flux) dw[j] -= flux_coef * wall[N];
        }
        thomas(o.a.data(), bw.data(), o.c.data(), dw.data(), f, N + 1, cp.data(), dp.data());
    }
    // src/FieldSolver.cpp:263-271
    void gradient(const Real* f, Real* df) const {
        const int N = g->N;
        for (int j = 0; j <= N; ++j) {
            if (j == 0) { df[0] = Real(0); continue; }
            if (j == N) { df[N] = g0[N] * f[N] + gm[N] * f[N - 1] + gp[N] * f[N - 2]; continue; }
            df[j] = gm[j] * f[j - 1] + g0[j] * f[j] + gp[j] * f[j + 1];
        }
    }
    // src/FieldSolver.cpp:273-281
    void gradient_odd(const Real* f, Real* df) const {
        const int N = g->N;
        for (int j = 0; j <= N; ++j) {
            if (j == 0) { df[0] = f[1] * ir[0]; continue; }
            if (j == N) { df[N] = g0[N] * f[N] + gm[N] * f[N - 1] + gp[N] * f[N - 2]; continue; }
            df[j] = gm[j] * f[j - 1] + g0[j] * f[j] + gp[j] * f[j + 1];
        }
    }
    // src/FieldSolver.cpp:283-289
    void over_r(const Real* f, Real* out) const {
        const int N = g->N;
        for (int j = 0; j <= N; ++j) out[j] = (j == 0 ? f[1] : f[j]) * ir[j];
    }
};

// ============================================================================ pushers (src/Pushers.hpp)
struct V3 { Real x, y, z; };
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(Real s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline Real dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// src/Pushers.hpp:36-39
inline Real gamma2_root(Real x, Real c) {
    const Real s = std::sqrt(x * x + c);
    return x >= Real(0) ? x + s : (s > -x ? c / (s - x) : Real(1));
}
// src/Pushers.hpp:44-52
inline V3 solve_implicit(V3 a, V3 b) {
    const Real b2 = dot(b, b), ab = dot(a, b);
    const Real g2 = gamma2_root(Real(0.5) * (Real(1) + dot(a, a) - b2), b2 + ab * ab);
    const Real g = std::sqrt(g2);
    const V3 axb = cross(a, b);
    const Real inv = Real(1) / (g2 + b2);
    return {inv * (g2 * a.x + g * axb.x + ab * b.x), inv * (g2 * a.y + g * axb.y + ab * b.y),
            inv * (g2 * a.z + g * axb.z + ab * b.z)};
}
// src/Pushers.hpp:54-61
inline V3 boris(V3 u, V3 E, V3 B, Real eps) {
    const V3 um = u + eps * E;
    const Real g = std::sqrt(Real(1) + dot(um, um));
    const V3 t = (eps / g) * B;
    const V3 s = (Real(2) / (Real(1) + dot(t, t))) * t;
    const V3 up = um + cross(um + cross(um, t), s);
    return up + eps * E;
}
// src/Pushers.hpp:63-68
inline V3 vay(V3 u, V3 E, V3 B, Real eps) {
    const Real g0 = std::sqrt(Real(1) + dot(u, u));
    const V3 up = u + (Real(2) * eps) * E + (eps / g0) * cross(u, B);
    return solve_implicit(up, eps * B);
}
// src/Pushers.hpp:70-80
inline V3 higuera_cary(V3 u, V3 E, V3 B, Real eps) {
    const V3 um = u + eps * E;
    const V3 tau = eps * B;
    const Real ustar = dot(um, tau);
    const Real sigma = Real(1) + dot(um, um) - dot(tau, tau);
    const Real gp = std::sqrt(Real(0.5) * (sigma + std::sqrt(sigma * sigma + Real(4) * (dot(tau, tau) + ustar * ustar))));
    const V3 t = (Real(1) / gp) * tau;
    const Real s = Real(1) / (Real(1) + dot(t, t));
    const V3 upl = s * (um + dot(um, t) * t + cross(um, t));
    return upl + eps * E + cross(upl, t);
}
// src/Pushers.hpp:83-86
inline V3 imp(V3 u, V3 E, V3 B, Real eps) {
    const V3 ubar = solve_implicit(u + eps * E, eps * B);
    return Real(2) * ubar - u;
}
// src/Pushers.hpp:90-95
inline V3 imp_rr(V3 u, V3 E, V3 B, Real eps, Real nu_dt) {
    const Real g0 = std::sqrt(Real(1) + dot(u, u));
    const Real f = Real(1) / (Real(1) + nu_dt);
    const V3 a = f * (u + (Real(2) * eps) * E + (eps / g0) * cross(u, B));
    return solve_implicit(a, (f * eps) * B);
}
inline V3 push_u(int mode, V3 u, V3 E, V3 B, Real eps) {
    switch (mode) {
        case 0: return vay(u, E, B, eps);
        case 1: return boris(u, E, B, eps);
        case 2: return higuera_cary(u, E, B, eps);
        case 3: return imp(u, E, B, eps);
        default: return imp_rr(u, E, B, eps, 0.0);
    }
}

// ============================================================================ Adams-Bashforth (src/PlasmaSpecies.cpp:14-48)
struct ABCoeffs { Real c[5] = {0, 0, 0, 0, 0}; int order = 1; };
static ABCoeffs make_ab(int order) {
    ABCoeffs ab;
    ab.order = order;
    switch (order) {
        case 1: ab.c[0] = 1; break;
        case 2: ab.c[0] = 1.5; ab.c[1] = -0.5; break;
        case 3: ab.c[0] = 23. / 12; ab.c[1] = -16. / 12; ab.c[2] = 5. / 12; break;
        case 4: ab.c[0] = 55. / 24; ab.c[1] = -59. / 24; ab.c[2] = 37. / 24; ab.c[3] = -9. / 24; break;
        case 5:
            ab.c[0] = 1901. / 720; ab.c[1] = -2774. / 720; ab.c[2] = 2616. / 720;
            ab.c[3] = -1274. / 720; ab.c[4] = 251. / 720; break;
        default: throw std::runtime_error("pusher.ab_order must be 1..5");
    }
    return ab;
}
static ABCoeffs make_ab_variable(int order, const double* t, double h) {
    ABCoeffs ab;
    ab.order = order;
    const double gx[3] = {-0.7745966692414834, 0.0, 0.7745966692414834};
    const double gw[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
    for (int m = 0; m < order; ++m) {
        double s = 0;
        for (int q = 0; q < 3; ++q) {
            const double tt = t[0] + 0.5 * h * (1.0 + gx[q]);
            double l = 1;
            for (int n = 0; n < order; ++n)
                if (n != m) l *= (tt - t[n]) / (t[m] - t[n]);
            s += 0.5 * gw[q] * l;
        }
        ab.c[m] = s;
    }
    return ab;
}


// ============================================================================ slice arrays (src/SliceData.hpp)
struct MScalar { std::vector<Real> a0, ar, ai; void init(int M) { a0.assign(M, 0); ar.assign(M, 0); ai.assign(M, 0); } };
struct MVector {
    std::vector<Real> r0, i0, r1, i1, r2, i2;
    void init(int M) { r0.assign(M, 0); i0.assign(M, 0); r1.assign(M, 0); i1.assign(M, 0); r2.assign(M, 0); i2.assign(M, 0); }
};
struct SliceSources {
    MScalar rhot, drho, jz, chi, rho, ne, ni;
    MVector jp, S;
    void init(int M) { rhot.init(M); drho.init(M); jz.init(M); chi.init(M); rho.init(M); ne.init(M); ni.init(M); jp.init(M); S.init(M); }
};
struct SliceFields {
    MScalar psi, ez, bz;
    MVector wp, bp, bprev;
    std::vector<Real> t1, t2, t3, t4, rhs;
    void init(int M) {
        psi.init(M); ez.init(M); bz.init(M); wp.init(M); bp.init(M); bprev.init(M);
        t1.assign(M, 0); t2.assign(M, 0); t3.assign(M, 0); t4.assign(M, 0); rhs.assign(M, 0);
    }
};
// src/SliceData.hpp:70-83
inline Real eval_scalar(Real f0, Real fr, Real fi, Real c, Real s, bool m1) {
    return m1 ? f0 + Real(2) * (fr * c - fi * s) : f0;
}
inline void eval_vector(Real r0, Real i0, Real r1, Real i1, Real r2, Real i2, Real c, Real s, bool m1, Real& vx, Real& vy) {
    vx = r1 * c - i1 * s;
    vy = r1 * s + i1 * c;
    if (m1) {
        const Real c2 = c * c - s * s, s2 = Real(2) * c * s;
        vx += r0 + r2 * c2 - i2 * s2;
        vy += i0 + r2 * s2 + i2 * c2;
    }
}

// field layouts (src/FieldStore.hpp:9-21, src/Simulation.hpp:46)
enum FieldComp : int {
    F_EP1R = 0, F_EP1I, F_BP1R, F_BP1I, F_EZ0, F_BZ0,
    F_EP0R, F_EP0I, F_EP2R, F_EP2I, F_BP0R, F_BP0I, F_BP2R, F_BP2I,
    F_EZ1R, F_EZ1I, F_BZ1R, F_BZ1I,
    F_NC0 = 6, F_NC1 = 18
};
enum BeamComp : int {
    B_RT0 = 0, B_JZ0, B_JP1R, B_JP1I, B_RHO0,
    B_RT1R, B_RT1I, B_JZ1R, B_JZ1I, B_JP0R, B_JP0I, B_JP2R, B_JP2I, B_RHO1R, B_RHO1I,
    B_NC0 = 5, B_NC1 = 15
};
enum DiagComp : int { D_PSI0 = 0, D_NE0, D_NI0, D_PSI1R, D_PSI1I, D_NE1R, D_NE1I, D_NI1R, D_NI1I, D_NC0 = 3, D_NC1 = 9 };

template <bool PAR> inline void add(Real& t, Real v) {
    if (PAR) {
#pragma omp atomic
        t += v;
    } else {
        t += v;
    }
}

// ============================================================================ particle geometry (src/PlasmaSpecies.cpp:52-89)
struct PGeo { Real r, c, s; int j; Real W[2], dW[2], Wor[2]; };
inline PGeo geometry(const Grid& g, Real x, Real y) {
    PGeo p;
    p.r = std::sqrt(x * x + y * y);
    if (p.r > Real(0)) { p.c = x / p.r; p.s = y / p.r; } else { p.c = Real(1); p.s = Real(0); }
    Real t;
    g.weights(p.r, p.j, t);
    const Real ih = g.inv_h[p.j];
    p.W[0] = Real(1) - t; p.W[1] = t;
    p.dW[0] = -ih; p.dW[1] = ih;
    const Real ir = p.r > Real(0) ? Real(1) / p.r : Real(0);
    p.Wor[0] = (p.j == 0) ? Real(0) : p.W[0] * ir;
    p.Wor[1] = (p.j == 0) ? ih : p.W[1] * ir;
    return p;
}
inline Real lin(const std::vector<Real>& f, const PGeo& p) { return p.W[0] * f[p.j] + p.W[1] * f[p.j + 1]; }

struct PFields { Real Wx, Wy, Bx, By, Bz, Ez, A, dA; };
inline PFields gather(const SliceFields& f, const PGeo& p, bool m1, bool needB) {
    PFields o;
    eval_vector(lin(f.wp.r0, p), lin(f.wp.i0, p), lin(f.wp.r1, p), lin(f.wp.i1, p), lin(f.wp.r2, p),
                lin(f.wp.i2, p), p.c, p.s, m1, o.Wx, o.Wy);
    o.Ez = eval_scalar(lin(f.ez.a0, p), lin(f.ez.ar, p), lin(f.ez.ai, p), p.c, p.s, m1);
    o.Bz = eval_scalar(lin(f.bz.a0, p), lin(f.bz.ar, p), lin(f.bz.ai, p), p.c, p.s, m1);
    if (needB)
        eval_vector(lin(f.bp.r0, p), lin(f.bp.i0, p), lin(f.bp.r1, p), lin(f.bp.i1, p), lin(f.bp.r2, p),
                    lin(f.bp.i2, p), p.c, p.s, m1, o.Bx, o.By);
    else
        o.Bx = o.By = Real(0);
    o.A = o.dA = Real(0);   // no laser envelope on this path
    return o;
}

// ============================================================================ plasma species (src/PlasmaSpecies.*)
struct Species {
    std::string name;
    const Grid* g = nullptr;
    bool m1 = false;
    Real q = -1, m = 1;
    bool mobile = true;
    int ppc = 4, ntheta = 1;
    Real uth = 0;             // constant thermal spread (cold plasma: 0)
    Real density = 1;         // <species>.density factor
    Real rmax = 1e300;        // plasma.rmax (uniform radial profile)
    Real delta_min = 1e-3, max_qsa = 35;
    ABCoeffs ab;
    long npush = 0;
    double xi_cur = 0;
    double xi_slot[5] = {0, 0, 0, 0, 0};
    int Np = 0, cap = 0;
    std::vector<Real> x, y, px, py, dl, w;
    std::vector<Real> hist;   // (order, 5, cap)
    long lost = 0;

    Real& H(int sl, int v, int i) { return hist[(static_cast<size_t>(sl) * 5 + v) * cap + i]; }

    // src/PlasmaSpecies.cpp:92-185 (keys of the path only)
    void init(const Config& cfg, const std::string& nm, const Grid& grid, bool mode1) {
        name = nm; g = &grid; m1 = mode1;
        q = cfg.get_double(nm + ".charge", -1.0);
        m = cfg.get_double(nm + ".mass", 1.0);
        mobile = cfg.get_bool(nm + ".mobile", true);
        ppc = cfg.get_int(nm + ".ppc", 4);
        ntheta = cfg.get_int(nm + ".ntheta", m1 ? 8 : 1);
        uth = cfg.get_double(nm + ".uth", 0.0);
        density = cfg.get_double(nm + ".density", 1.0);
        rmax = cfg.get_double("plasma.rmax", 1e300);
        const std::string rad = cfg.get_string("plasma.radial", "uniform");
        if (rad != "uniform") throw std::runtime_error("oracle: only plasma.radial = uniform");
        delta_min = cfg.get_double("pusher.delta_min", 1e-3);
        max_qsa = cfg.get_double("pusher.max_qsa_factor", 35.0);
        ab = make_ab(cfg.get_int("pusher.ab_order", 3));
        if (ppc < 1) throw std::runtime_error(nm + ".ppc must be >= 1");
        if (ntheta < 1) throw std::runtime_error(nm + ".ntheta must be >= 1");
        if (m1 && ntheta < 4) throw std::runtime_error(nm + ".ntheta must be >= 4 with azimuthal mode 1");
        cap = std::max(grid.N * ppc * ntheta, 1);
        Np = cap;
        x.assign(cap, 0); y.assign(cap, 0); px.assign(cap, 0); py.assign(cap, 0); dl.assign(cap, 0); w.assign(cap, 0);
        hist.assign(static_cast<size_t>(ab.order) * 5 * cap, 0);
    }

    // DensityProfile::radial for the uniform profile (src/Profiles.cpp:52-57)
    Real radial(Real r) const { return r > rmax ? 0.0 : 1.0; }

    // src/PlasmaSpecies.cpp:202-295 (uniform profile, constant uth, no ionization)
    void load(Real z, unsigned long long seed) {
        (void)z;
        std::mt19937_64 rng(seed);
        std::normal_distribution<double> gauss(0.0, 1.0);
        const double fz = 1.0 * density;   // longitudinal(z) = 1 without a z table
        int i = 0;
        const Grid& G = *g;
        for (int j = 0; j < G.N; ++j) {
            const double h = G.h[j];
            for (int mm = 0; mm < ppc; ++mm) {
                const double ra = G.r[j] + mm * h / ppc, rb = G.r[j] + (mm + 1) * h / ppc;
                const double r = (2.0 / 3.0) * (rb * rb * rb - ra * ra * ra) / (rb * rb - ra * ra);
                const double vring = fz * PI * (rb * rb - ra * ra);
                const double wring = radial(r) * vring;
                const double off = ((j * ppc + mm) % 2) * 0.5;
                for (int l = 0; l < ntheta; ++l, ++i) {
                    const double th = 2 * PI * (l + off) / ntheta;
                    double ppx = 0, ppy = 0, ppz = 0;
                    const double ut = uth;
                    if (ut > 0) { ppx = ut * gauss(rng); ppy = ut * gauss(rng); ppz = ut * gauss(rng); }
                    ppx += 0; ppy += 0; ppz += 0;   // u_drift = 0
                    const double gam = std::sqrt(1 + ppx * ppx + ppy * ppy + ppz * ppz);
                    const double X = (ntheta == 1) ? r : r * std::cos(th), Y = (ntheta == 1) ? 0.0 : r * std::sin(th);
                    const double W = wring / ntheta;
                    if (W == 0) { --i; continue; }
                    x[i] = X; y[i] = Y; px[i] = ppx; py[i] = ppy; dl[i] = gam - ppz; w[i] = W;
                }
            }
        }
        std::fill(hist.begin(), hist.end(), 0.0);
        lost = 0;
        Np = i;
        npush = 0;
        xi_cur = 0;
    }

    // src/PlasmaSpecies.cpp:409-478
    template <bool PAR> void deposit(SliceSources& s) const {
        const Grid& G = *g;
        const Real q0 = q, qm0 = q / m;
        const bool neg = q < 0;
        const Real chif = Real(1);
        MScalar& dens = neg ? s.ne : s.ni;
        SliceSources* sp = &s;
#pragma omp parallel for if (PAR)
        for (int i = 0; i < Np; ++i) {
            const Real wi = w[i];
            if (wi == Real(0)) continue;
            const Real zi = Real(1);
            const Real qq = q0 * zi, qm = qm0 * zi;
            const Real PX = px[i], PY = py[i], D = dl[i];
            const PGeo p = geometry(G, x[i], y[i]);
            const Real gam = (Real(1) + PX * PX + PY * PY + D * D) / (Real(2) * D);
            const Real invd = Real(1) / D;
            const Real pz = gam - D;
            const Real ux = PX * invd, uy = PY * invd;
            const Real ur = p.c * ux + p.s * uy, uth_ = p.c * uy - p.s * ux;
            const Real wq = wi * qq;
            for (int cc = 0; cc < 2; ++cc) {
                const int jj = p.j + cc;
                const Real iv = G.inv_V[jj];
                const Real a = wq * p.W[cc] * iv;
                const Real dr = wq * ur * p.dW[cc] * iv;
                const Real vjz = a * pz * invd, vchi = chif * a * qm * invd, vrho = a * gam * invd;
                const Real vn = wi * p.W[cc] * iv * gam * invd;
                add<PAR>(sp->rhot.a0[jj], a);
                add<PAR>(sp->drho.a0[jj], dr);
                add<PAR>(sp->jz.a0[jj], vjz);
                add<PAR>(sp->chi.a0[jj], vchi);
                add<PAR>(sp->rho.a0[jj], vrho);
                add<PAR>(dens.a0[jj], vn);
                add<PAR>(sp->jp.r1[jj], a * ur);
                add<PAR>(sp->jp.i1[jj], a * uth_);
                if (m1) {
                    const Real c = p.c, sn = p.s;
                    add<PAR>(sp->rhot.ar[jj], a * c);
                    add<PAR>(sp->rhot.ai[jj], -a * sn);
                    const Real th = wq * uth_ * p.Wor[cc] * iv;
                    add<PAR>(sp->drho.ar[jj], dr * c - th * sn);
                    add<PAR>(sp->drho.ai[jj], -dr * sn - th * c);
                    add<PAR>(sp->jz.ar[jj], vjz * c);
                    add<PAR>(sp->jz.ai[jj], -vjz * sn);
                    add<PAR>(sp->chi.ar[jj], vchi * c);
                    add<PAR>(sp->chi.ai[jj], -vchi * sn);
                    add<PAR>(sp->rho.ar[jj], vrho * c);
                    add<PAR>(sp->rho.ai[jj], -vrho * sn);
                    add<PAR>(dens.ar[jj], vn * c);
                    add<PAR>(dens.ai[jj], -vn * sn);
                    add<PAR>(sp->jp.r0[jj], a * ux);
                    add<PAR>(sp->jp.i0[jj], a * uy);
                    const Real c2 = c * c - sn * sn, s2 = Real(2) * c * sn;
                    add<PAR>(sp->jp.r2[jj], a * (ux * c2 + uy * s2));
                    add<PAR>(sp->jp.i2[jj], a * (uy * c2 - ux * s2));
                }
            }
        }
    }

    // src/PlasmaSpecies.cpp:480-506
    template <bool PAR> void deposit_rho(MScalar& rhot, MScalar& rho) const {
        const Grid& G = *g;
        const Real q0 = q;
#pragma omp parallel for if (PAR)
        for (int i = 0; i < Np; ++i) {
            const Real wi = w[i];
            if (wi == Real(0)) continue;
            const Real qq = q0 * Real(1);
            if (qq == Real(0)) continue;
            const PGeo p = geometry(G, x[i], y[i]);
            for (int cc = 0; cc < 2; ++cc) {
                const int jj = p.j + cc;
                const Real a = wi * qq * p.W[cc] * G.inv_V[jj];
                add<PAR>(rhot.a0[jj], a);
                add<PAR>(rho.a0[jj], a);
                if (m1) {
                    add<PAR>(rhot.ar[jj], a * p.c);
                    add<PAR>(rhot.ai[jj], -a * p.s);
                    add<PAR>(rho.ar[jj], a * p.c);
                    add<PAR>(rho.ai[jj], -a * p.s);
                }
            }
        }
    }

    // src/PlasmaSpecies.cpp:508-559
    template <bool PAR> void deposit_S(const SliceFields& f, MVector& S) const {
        const Grid& G = *g;
        const Real q0 = q, qm0 = q / m;
        MVector* Sp = &S;
#pragma omp parallel for if (PAR)
        for (int i = 0; i < Np; ++i) {
            const Real wi = w[i];
            if (wi == Real(0)) continue;
            const Real zi = Real(1);
            const Real qq = q0 * zi, qm = qm0 * zi;
            const Real PX = px[i], PY = py[i], D = dl[i];
            const PGeo p = geometry(G, x[i], y[i]);
            const PFields F = gather(f, p, m1, false);
            const Real gam = (Real(1) + PX * PX + PY * PY + D * D) / (Real(2) * D);
            const Real invd = Real(1) / D;
            const Real fp = -Real(0.5) * qm * qm * F.dA * invd;
            const Real ux = PX * invd, uy = PY * invd;
            const Real ur = p.c * ux + p.s * uy, uth_ = p.c * uy - p.s * ux;
            const Real dD = qm * ((PX * F.Wx + PY * F.Wy) * invd - F.Ez);
            const Real ax = (qm * (gam * F.Wx + PY * F.Bz) * invd + fp * p.c) * invd - PX * dD * invd * invd;
            const Real ay = (qm * (gam * F.Wy - PX * F.Bz) * invd + fp * p.s) * invd - PY * dD * invd * invd;
            const Real wq = wi * qq;
            auto add_mode = [&](int k, std::vector<Real>& Sr, std::vector<Real>& Si, Real ck, Real sk) {
                for (int cc = 0; cc < 2; ++cc) {
                    const int jj = p.j + cc;
                    const Real iv = G.inv_V[jj] * wq;
                    Real Ar = ax * p.W[cc] + ux * ur * p.dW[cc];
                    Real Ai = ay * p.W[cc] + uy * ur * p.dW[cc];
                    if (k != 0) {
                        const Real f2 = k * uth_ * p.Wor[cc];
                        Ar += f2 * uy;
                        Ai -= f2 * ux;
                    }
                    add<PAR>(Sr[jj], iv * (Ar * ck + Ai * sk));
                    add<PAR>(Si[jj], iv * (Ai * ck - Ar * sk));
                }
            };
            add_mode(1, Sp->r1, Sp->i1, p.c, p.s);
            if (m1) {
                add_mode(0, Sp->r0, Sp->i0, Real(1), Real(0));
                add_mode(2, Sp->r2, Sp->i2, p.c * p.c - p.s * p.s, Real(2) * p.c * p.s);
            }
        }
    }

    // history bookkeeping of push (src/PlasmaSpecies.cpp:561-581): slot, first flag and coefficients
    void push_setup(Real dxi, int& slot, bool& first, ABCoeffs& abv) {
        const int ordh = ab.order;
        const int k = static_cast<int>(npush % ordh);
        const bool first0 = (npush == 0);
        if (first0)
            for (int mm = 0; mm < ordh; ++mm) xi_slot[(k - mm + ordh) % ordh] = xi_cur - mm * static_cast<double>(dxi);
        else
            xi_slot[k] = xi_cur;
        double tn[5];
        bool uniform = true;
        for (int mm = 0; mm < ordh; ++mm) {
            tn[mm] = xi_slot[(k - mm + ordh) % ordh];
            if (mm > 0 && std::abs((tn[mm - 1] - tn[mm]) - static_cast<double>(dxi)) > 1e-9 * dxi) uniform = false;
        }
        abv = uniform ? ab : make_ab_variable(ordh, tn, dxi);
        ++npush;
        xi_cur += dxi;
        slot = k;
        first = first0;
    }

    // src/PlasmaSpecies.cpp:561-668
    template <bool PAR> void push(const SliceFields& f, Real dxi) {
        if (!mobile) return;
        int slot; bool first; ABCoeffs abv;
        push_setup(dxi, slot, first, abv);
        const Grid& G = *g;
        const Real qm0 = q / m;
        const Real dmin = delta_min, qsamax = max_qsa;
        const int ord = abv.order;
        const Real R = G.R;
        long nlost = 0;
#pragma omp parallel for if (PAR) reduction(+ : nlost)
        for (int i = 0; i < Np; ++i) {
            const Real wi = w[i];
            if (wi == Real(0)) continue;
            const Real X = x[i], Y = y[i], PX = px[i], PY = py[i], D = dl[i];
            const Real zi = Real(1);
            const Real qm = qm0 * zi;
            const PGeo p = geometry(G, X, Y);
            const PFields F = gather(f, p, m1, true);
            const Real gam = (Real(1) + PX * PX + PY * PY + D * D) / (Real(2) * D);
            const Real invd = Real(1) / D;
            const Real fp = -Real(0.5) * qm * qm * F.dA * invd;
            Real fv[5];
            fv[0] = PX * invd;
            fv[1] = PY * invd;
            fv[2] = qm * (gam * F.Wx * invd + F.By + PY * F.Bz * invd) + fp * p.c;
            fv[3] = qm * (gam * F.Wy * invd - F.Bx - PX * F.Bz * invd) + fp * p.s;
            fv[4] = qm * ((PX * F.Wx + PY * F.Wy) * invd - F.Ez);
            if (first) {
                for (int sl = 0; sl < ord; ++sl)
                    for (int v = 0; v < 5; ++v) H(sl, v, i) = fv[v];
            } else {
                for (int v = 0; v < 5; ++v) H(slot, v, i) = fv[v];
            }
            Real sum[5] = {0, 0, 0, 0, 0};
            for (int mm = 0; mm < ord; ++mm) {
                const int sl = (slot - mm + ord) % ord;
                for (int v = 0; v < 5; ++v) sum[v] += abv.c[mm] * H(sl, v, i);
            }
            Real xn = X + dxi * sum[0], yn = Y + dxi * sum[1];
            Real pxn = PX + dxi * sum[2], pyn = PY + dxi * sum[3];
            const Real dn = D + dxi * sum[4];
            const Real gn = (Real(1) + pxn * pxn + pyn * pyn + dn * dn) / (Real(2) * dn);
            if (!(dn > dmin) || gn > qsamax * dn) {
                w[i] = Real(0);
                ++nlost;
                continue;
            }
            const Real rn = std::sqrt(xn * xn + yn * yn);
            if (rn > R) {
                const Real nx = xn / rn, ny = yn / rn;
                const Real rr = Real(2) * R - rn;
                if (rr < Real(0)) { w[i] = Real(0); ++nlost; continue; }
                xn = nx * rr; yn = ny * rr;
                const Real pn = pxn * nx + pyn * ny;
                pxn -= Real(2) * pn * nx; pyn -= Real(2) * pn * ny;
                for (int sl = 0; sl < ord; ++sl) {
                    const Real aa = H(sl, 0, i) * nx + H(sl, 1, i) * ny;
                    H(sl, 0, i) -= Real(2) * aa * nx; H(sl, 1, i) -= Real(2) * aa * ny;
                    const Real bb = H(sl, 2, i) * nx + H(sl, 3, i) * ny;
                    H(sl, 2, i) -= Real(2) * bb * nx; H(sl, 3, i) -= Real(2) * bb * ny;
                }
            }
            x[i] = xn; y[i] = yn; px[i] = pxn; py[i] = pyn; dl[i] = dn;
        }
        lost += nlost;
    }
};


// ============================================================================ beams (src/Beam.*)
struct BeamGrid { Real xi_min = 0, dxi = 0; int nxi = 0, k0 = 0, nloc = 0; bool ngp = false; };

// Beam::deposit as a free function over SoA arrays (src/Beam.cpp:725-785); src(K, M, nc) LayoutRight
template <bool PAR>
static void beam_deposit(const Grid& g, const BeamGrid& bg, bool m1, Real q, int Np, const Real* x, const Real* y,
                         const Real* px, const Real* py, const Real* pz, const Real* xi, const Real* w, Real* src,
                         int nc) {
    const Real xmin = bg.xi_min, idxi = Real(1) / bg.dxi;
    const int k0 = bg.k0, nloc = bg.nloc;
    const bool ngp = bg.ngp;
    const Real R = g.R;
    const int M = g.N + 1;
    auto S = [&](int kk, int jj, int comp) -> Real& { return src[(static_cast<size_t>(kk) * M + jj) * nc + comp]; };
#pragma omp parallel for if (PAR)
    for (int i = 0; i < Np; ++i) {
        const Real wi = w[i];
        if (wi == Real(0)) continue;
        const Real X = x[i], Y = y[i];
        const Real r = std::sqrt(X * X + Y * Y);
        if (r > R) continue;
        const Real c = r > Real(0) ? X / r : Real(1), s = r > Real(0) ? Y / r : Real(0);
        const Real sx = (xi[i] - xmin) * idxi;
        const int k = ngp ? static_cast<int>(std::floor(sx + Real(0.5))) : static_cast<int>(std::floor(sx));
        if (k < k0 - 1 || k >= k0 + nloc) continue;
        const Real tk = ngp ? Real(0) : sx - k;
        const Real p2 = px[i] * px[i] + py[i] * py[i];
        const Real gam = std::sqrt(Real(1) + p2 + pz[i] * pz[i]);
        const Real omvz = (Real(1) + p2) / (gam * (gam + pz[i]));
        const Real vz = pz[i] / gam, vx = px[i] / gam, vy = py[i] / gam;
        const Real vr = c * vx + s * vy, vt = c * vy - s * vx;
        int j; Real t;
        g.weights(r, j, t);
        const Real Wr[2] = {Real(1) - t, t};
        const Real Wk[2] = {Real(1) - tk, tk};
        for (int a = 0; a < (ngp ? 1 : 2); ++a) {
            const int kk = k + a - k0;
            if (kk < 0 || kk >= nloc) continue;
            for (int b = 0; b < 2; ++b) {
                const int jj = j + b;
                const Real d = wi * q * Wk[a] * Wr[b] * g.inv_V[jj] * idxi;
                add<PAR>(S(kk, jj, B_RT0), d * omvz);
                add<PAR>(S(kk, jj, B_JZ0), d * vz);
                add<PAR>(S(kk, jj, B_JP1R), d * vr);
                add<PAR>(S(kk, jj, B_JP1I), d * vt);
                add<PAR>(S(kk, jj, B_RHO0), d);
                if (m1) {
                    add<PAR>(S(kk, jj, B_RT1R), d * omvz * c);
                    add<PAR>(S(kk, jj, B_RT1I), -d * omvz * s);
                    add<PAR>(S(kk, jj, B_JZ1R), d * vz * c);
                    add<PAR>(S(kk, jj, B_JZ1I), -d * vz * s);
                    add<PAR>(S(kk, jj, B_RHO1R), d * c);
                    add<PAR>(S(kk, jj, B_RHO1I), -d * s);
                    add<PAR>(S(kk, jj, B_JP0R), d * vx);
                    add<PAR>(S(kk, jj, B_JP0I), d * vy);
                    const Real c2 = c * c - s * s, s2 = Real(2) * c * s;
                    add<PAR>(S(kk, jj, B_JP2R), d * (vx * c2 + vy * s2));
                    add<PAR>(S(kk, jj, B_JP2I), d * (vy * c2 - vx * s2));
                }
            }
        }
    }
}

// e^{-z} I0(z) and e^{-z} I1(z) for z >= 0 by the polynomial approximations of Abramowitz & Stegun,
// Handbook of Mathematical Functions, 9.8.1-9.8.4 (|relative error| < 2e-7).  The reference uses the same
// published approximation for its analytic beam (src/Beam.cpp:217-233); the coefficients are A&S's table.
static void scaled_bessel_i0_i1(Real z, Real& i0, Real& i1) {
    if (z < Real(3.75)) {
        const Real u = (z / Real(3.75)) * (z / Real(3.75));
        // A&S 9.8.1 and 9.8.3 (polynomials in (z/3.75)^2), Horner order
        const Real p0 = 1 + u * (3.5156229 + u * (3.0899424 + u * (1.2067492 + u * (0.2659732 + u * (0.0360768 + u * 0.0045813)))));
        const Real p1 = z * (0.5 + u * (0.87890594 + u * (0.51498869 + u * (0.15084934 + u * (0.02658733 + u * (0.00301532 + u * 0.00032411))))));
        const Real e = std::exp(-z);
        i0 = p0 * e;
        i1 = p1 * e;
        return;
    }
    // A&S 9.8.2 and 9.8.4 (asymptotic polynomials in 3.75/z), already scaled by e^{-z}
    const Real v = Real(3.75) / z;
    const Real rs = Real(1) / std::sqrt(z);
    i0 = rs * (0.39894228 + v * (0.01328592 + v * (0.00225319 + v * (-0.00157565 + v * (0.00916281 + v * (-0.02057706 + v * (0.02635537 + v * (-0.01647633 + v * 0.00392377))))))));
    i1 = rs * (0.39894228 + v * (-0.03988024 + v * (-0.00362018 + v * (0.00163801 + v * (-0.01031555 + v * (0.02282967 + v * (-0.02895312 + v * (0.01787654 - v * 0.00420059))))))));
}

// Rigid analytic beam (deck key <beam>.analytic = 1), Beam::deposit_analytic (src/Beam.cpp:235-281): a Gaussian
// of peak density n0 and rms radius sr whose centroid c(xi) = (x0, y0) + (xs, ys)(xi - xi0) sits at distance d
// from the axis.  Its azimuthal modes at radius r follow from the generating function of I_n:
// exp(-(r^2 + d^2 - 2 r d cos(th - phi)) / 2 sr^2) = e^{-(r-d)^2/2sr^2} [e^{-z} I0(z) + 2 e^{-z} I1(z) cos(th - phi) + ...],
// z = r d / sr^2; mode 1 carries the phase e^{-i phi}.  The source is added at every node of every slice (no
// particles).  Oracle-only: it drives the m = 1 validation runs of validation/run_validation.sh section 6.
struct AnalyticBeam {
    Real n0 = 0, sr = 0, xi0 = 0, sxi = 0, len = 0, gamma = 0, cut = -1e300, x0 = 0, y0 = 0, xs = 0, ys = 0;
    bool flat = false;
};
static void beam_deposit_analytic(const Grid& g, const BeamGrid& bg, bool m1, Real q, const AnalyticBeam& a, Real* src,
                                  int nc) {
    const int M = g.N + 1;
    const Real s2 = a.sr * a.sr;
    const Real vz = std::sqrt(Real(1) - Real(1) / (a.gamma * a.gamma));
    const Real omvz = Real(1) / (a.gamma * a.gamma * (Real(1) + vz));
    for (int k = 0; k < bg.nloc; ++k) {
        const Real xi = bg.xi_min + (bg.k0 + k) * bg.dxi;
        if (xi < a.cut) continue;
        const Real fxi = a.flat ? ((std::fabs(xi - a.xi0) <= Real(0.5) * a.len) ? Real(1) : Real(0))
                                : std::exp(-Real(0.5) * (xi - a.xi0) * (xi - a.xi0) / (a.sxi * a.sxi));
        if (fxi == Real(0)) continue;
        const Real cx = a.x0 + a.xs * (xi - a.xi0), cy = a.y0 + a.ys * (xi - a.xi0);
        const Real d = std::sqrt(cx * cx + cy * cy);
        for (int j = 0; j < M; ++j) {
            Real* S = src + (static_cast<size_t>(k) * M + j) * nc;
            const Real r = g.r[j];
            Real i0, i1;
            scaled_bessel_i0_i1(r * d / s2, i0, i1);
            const Real env = q * a.n0 * fxi * std::exp(-Real(0.5) * (r - d) * (r - d) / s2);
            const Real m0 = env * i0;
            S[B_RHO0] += m0;
            S[B_RT0] += m0 * omvz;
            S[B_JZ0] += m0 * vz;
            if (m1 && d > Real(0)) {
                const Real f1r = env * i1 * (cx / d), f1i = -env * i1 * (cy / d);
                S[B_RHO1R] += f1r;
                S[B_RHO1I] += f1i;
                S[B_RT1R] += f1r * omvz;
                S[B_RT1I] += f1i * omvz;
                S[B_JZ1R] += f1r * vz;
                S[B_JZ1I] += f1i * vz;
            }
        }
    }
}

// Beam::advance without laser, spin, radiation reaction or scattering (src/Beam.cpp:948-1131)
template <bool PAR>
static void beam_advance(const Grid& g, const BeamGrid& bg, bool m1, Real qm, int mode, int i0, int Np, Real* x,
                         Real* y, Real* px, Real* py, Real* pz, Real* xi, Real* w, const Real* fld, int nc,
                         Real dt_kick, Real dt, bool move) {
    const Real xmin = bg.xi_min, idxi = Real(1) / bg.dxi;
    const Real xmax = bg.xi_min + (bg.nxi - 1) * bg.dxi;
    const int k0 = bg.k0, nloc = bg.nloc;
    const bool ngp = bg.ngp;
    const Real R = g.R;
    const int M = g.N + 1;
    auto F = [&](int kk, int jj, int comp) { return fld[(static_cast<size_t>(kk) * M + jj) * nc + comp]; };
#pragma omp parallel for if (PAR)
    for (int i = i0; i < Np; ++i) {
        if (w[i] == Real(0)) continue;
        const Real X = x[i], Y = y[i];
        const Real r = std::sqrt(X * X + Y * Y);
        const Real c = r > Real(0) ? X / r : Real(1), sn = r > Real(0) ? Y / r : Real(0);
        Real s = (xi[i] - xmin) * idxi - k0;
        int k;
        Real tk;
        if (ngp) {
            k = static_cast<int>(std::floor(s + Real(0.5)));
            k = k < 0 ? 0 : (k > nloc - 1 ? nloc - 1 : k);
            tk = Real(0);
        } else {
            k = static_cast<int>(std::floor(s));
            if (k < 0) { k = 0; s = 0; }
            if (k > nloc - 2) { k = nloc - 2; s = nloc - 1; }
            tk = s - k;
        }
        const int k1 = ngp ? k : k + 1;
        int j; Real t;
        g.weights(r, j, t);
        auto bil = [&](int comp) {
            return (Real(1) - tk) * ((Real(1) - t) * F(k, j, comp) + t * F(k, j + 1, comp)) +
                   tk * ((Real(1) - t) * F(k1, j, comp) + t * F(k1, j + 1, comp));
        };
        Real Ex, Ey, Bx, By, Ez, Bz;
        if (m1) {
            eval_vector(bil(F_EP0R), bil(F_EP0I), bil(F_EP1R), bil(F_EP1I), bil(F_EP2R), bil(F_EP2I), c, sn, true, Ex, Ey);
            eval_vector(bil(F_BP0R), bil(F_BP0I), bil(F_BP1R), bil(F_BP1I), bil(F_BP2R), bil(F_BP2I), c, sn, true, Bx, By);
            Ez = eval_scalar(bil(F_EZ0), bil(F_EZ1R), bil(F_EZ1I), c, sn, true);
            Bz = eval_scalar(bil(F_BZ0), bil(F_BZ1R), bil(F_BZ1I), c, sn, true);
        } else {
            eval_vector(Real(0), Real(0), bil(F_EP1R), bil(F_EP1I), Real(0), Real(0), c, sn, false, Ex, Ey);
            eval_vector(Real(0), Real(0), bil(F_BP1R), bil(F_BP1I), Real(0), Real(0), c, sn, false, Bx, By);
            Ez = bil(F_EZ0);
            Bz = bil(F_BZ0);
        }
        const V3 E{Ex, Ey, Ez}, B{Bx, By, Bz};
        V3 u{px[i], py[i], pz[i]};
        const Real h = Real(0.5) * qm * dt_kick;
        const V3 Fp{Real(0), Real(0), Real(0)};   // no laser
        if (!move) {
            const Real g0 = std::sqrt(Real(1) + dot(u, u));
            u = u + (Real(2) * h) * (E + cross((Real(1) / g0) * u, B)) + dt_kick * Fp;
            px[i] = u.x; py[i] = u.y; pz[i] = u.z;
            continue;
        }
        u = u + (Real(0.5) * dt_kick) * Fp;
        u = push_u(mode, u, E, B, h);
        u = u + (Real(0.5) * dt_kick) * Fp;
        const Real p2 = u.x * u.x + u.y * u.y;
        const Real gam = std::sqrt(Real(1) + p2 + u.z * u.z);
        const Real omvz = (Real(1) + p2) / (gam * (gam + u.z));
        const Real xn = X + u.x / gam * dt, yn = Y + u.y / gam * dt;
        x[i] = xn; y[i] = yn;
        px[i] = u.x; py[i] = u.y; pz[i] = u.z;
        xi[i] += omvz * dt;
        if (xn * xn + yn * yn >= R * R || xi[i] < xmin || xi[i] > xmax) w[i] = Real(0);
    }
}

struct BeamDiag {
    double t = 0, npart = 0, charge = 0, gamma_mean = 0, gamma_rms = 0, xi_mean = 0, xi_rms = 0;
    double r_rms = 0, emit_nx = 0, emit_ny = 0, x_mean = 0, y_mean = 0, sigma_x = 0, sigma_y = 0;
    long alive = 0;
};

struct Beam {
    std::string name;
    bool m1 = false;
    Real q = -1, m = 1;
    bool rigid = false;
    bool analytic = false;
    AnalyticBeam an;
    int mode = 0;
    bool started = false;
    double gref = 0, xiref = 0;
    int Np = 0;
    std::vector<Real> x, y, px, py, pz, xi, w;
    // the piecewise-linear weight profile (empty without <beam>.pwl_weights): values, flat-top ends in user xi,
    // the macro weight at profile value 1, and every particle's xi as sampled (the adjoint's weight derivative)
    std::vector<double> pwl;
    double pwl_u0 = 0, pwl_u1 = 0, pwl_wmac = 0;
    std::vector<double> xi_init;

    // Beam ctor for profile = gaussian (src/Beam.cpp:53-207) and init_gaussian (src/Beam.cpp:483-570)
    void init(const Config& cfg, const std::string& nm, bool mode1) {
        name = nm; m1 = mode1;
        q = cfg.get_double(nm + ".charge", -1.0);
        m = cfg.get_double(nm + ".mass", 1.0);
        rigid = cfg.get_bool(nm + ".rigid", false);
        const std::string pm = cfg.get_string(nm + ".pusher", cfg.get_string("pusher.beam", "vay"));
        if (pm == "vay") mode = 0;
        else if (pm == "boris") mode = 1;
        else if (pm == "hc") mode = 2;
        else if (pm == "imp") mode = 3;
        else throw std::runtime_error("pusher.beam must be vay|boris|hc|imp");
        const std::string p = nm;
        const std::string lprof = cfg.get_string(p + ".profile", "gaussian");
        if (lprof != "gaussian" && lprof != "flattop") throw std::runtime_error("oracle: beam profile gaussian|flattop");
        analytic = cfg.get_bool(p + ".analytic", false);
        if (analytic) {   // rigid, no macro-particles (src/Beam.cpp:183-200)
            rigid = true;
            an.n0 = cfg.get_double(p + ".n0");
            an.sr = cfg.get_double(p + ".sigma_r");
            an.xi0 = xi_from_user(cfg.get_double(p + ".xi0"));
            an.gamma = cfg.get_double(p + ".gamma");
            an.flat = (lprof == "flattop");
            if (an.flat) an.len = cfg.get_double(p + ".length");
            else an.sxi = cfg.get_double(p + ".sigma_xi");
            if (cfg.has(p + ".xi_cut_head")) an.cut = xi_from_user(cfg.get_double(p + ".xi_cut_head"));
            an.x0 = cfg.get_double(p + ".x0", 0.0);
            an.y0 = cfg.get_double(p + ".y0", 0.0);
            an.xs = -cfg.get_double(p + ".x_slope", 0.0);
            an.ys = -cfg.get_double(p + ".y_slope", 0.0);
            if (!m1 && (an.x0 != 0 || an.y0 != 0 || an.xs != 0 || an.ys != 0))
                throw std::runtime_error(p + ": transverse offsets/tilts need azimuthal mode 1 (modes = 1)");
            set_particles({});
            gref = 0; xiref = 0;
            return;
        }
        const double n0 = cfg.get_double(p + ".n0");
        const double sr = cfg.get_double(p + ".sigma_r");
        const double xi0 = xi_from_user(cfg.get_double(p + ".xi0"));
        const double gamma0 = cfg.get_double(p + ".gamma");
        const double espread = cfg.get_double(p + ".espread", 0.0);
        const double emit = cfg.get_double(p + ".emit_n", 0.0);
        const double cut = cfg.get_double(p + ".cut", 3.0);
        const double cx0 = cfg.get_double(p + ".x0", 0.0), cy0 = cfg.get_double(p + ".y0", 0.0);
        const double cxs = -cfg.get_double(p + ".x_slope", 0.0), cys = -cfg.get_double(p + ".y_slope", 0.0);
        const double xp0 = cfg.get_double(p + ".xp0", 0.0), yp0 = cfg.get_double(p + ".yp0", 0.0);
        if (!m1 && (cx0 != 0 || cy0 != 0 || cxs != 0 || cys != 0 || xp0 != 0 || yp0 != 0))
            throw std::runtime_error(p + ": transverse offsets/tilts need azimuthal mode 1 (modes = 1)");
        Np = cfg.get_int(p + ".nparticles", 100000);
        const int nrot = cfg.get_int(p + ".nsym", 1);
        const bool mirror = cfg.get_bool(p + ".mirror_y", false);
        const int nsym = nrot * (mirror ? 2 : 1);
        if (nrot < 1 || Np % nsym != 0) throw std::runtime_error(p + ".nparticles must be a multiple of nsym");
        const unsigned long long seed = cfg.get_int(p + ".seed", 12345);
        const double xcut = cfg.has(p + ".xi_cut_head") ? xi_from_user(cfg.get_double(p + ".xi_cut_head")) : -1e300;
        double sxi = 0, len = 0;
        if (lprof == "gaussian") sxi = cfg.get_double(p + ".sigma_xi");
        else len = cfg.get_double(p + ".length");
        const double Pr = 1.0 - std::exp(-0.5 * cut * cut);
        auto Phi = [](double xx) { return 0.5 * (1.0 + std::erf(xx / std::sqrt(2.0))); };
        double wtot;
        if (lprof == "gaussian") {
            const double lo = std::max(-cut, (xcut - xi0) / sxi);
            if (lo >= cut) throw std::runtime_error(p + ": xi_cut_head removes the whole beam");
            wtot = n0 * libm_pow(2 * PI, 1.5) * sr * sr * sxi * Pr * (Phi(cut) - Phi(lo));
        } else {
            const double lo = std::max(xi0 - 0.5 * len, xcut), hi = xi0 + 0.5 * len;
            if (lo >= hi) throw std::runtime_error(p + ": xi_cut_head removes the whole beam");
            wtot = n0 * 2 * PI * sr * sr * Pr * (hi - lo);
        }
        const double wmac = wtot / Np;
        // build-defined piecewise-linear weight profile of a flat-top beam (the differentiable workload's
        // witness; not in the reference): weight = wmac * profile(user xi), the profile's values at the
        // centres of equal segments of the flat top, tail first, linear between centres, flat beyond them
        std::vector<double> pwl;
        for (const auto& tok : cfg.get_list(p + ".pwl_weights")) pwl.push_back(cfg.number(tok, p + ".pwl_weights"));
        if (!pwl.empty() && (lprof != "flattop" || nsym != 1)) throw std::runtime_error(p + ".pwl_weights: flattop, nsym 1");
        const double pu0 = xi_to_user(xi0 + 0.5 * len), pu1 = xi_to_user(xi0 - 0.5 * len);
        auto profile = [&](double u) {
            const int S = static_cast<int>(pwl.size());
            const double tt = (u - pu0) / (pu1 - pu0) * S - 0.5;
            if (!(tt > 0)) return pwl[0];
            if (tt >= S - 1) return pwl[S - 1];
            const int sg = static_cast<int>(tt);
            const double f = tt - sg;
            return pwl[sg] + (pwl[sg + 1] - pwl[sg]) * f;
        };
        std::mt19937_64 rng(seed);
        std::normal_distribution<double> N01(0.0, 1.0);
        std::uniform_real_distribution<double> U01(0.0, 1.0);
        const double spx = (sr > 0) ? emit / sr : 0.0;
        const int Ntot = Np;
        double aw = 0, awg = 0, awxi = 0;
        std::vector<double> local;
        local.reserve(static_cast<size_t>(Ntot) * 7);
        for (int i0 = 0; i0 < Ntot; i0 += nsym) {
            double xx, yy;
            do { xx = sr * N01(rng); yy = sr * N01(rng); } while (xx * xx + yy * yy > cut * cut * sr * sr);
            double xv;
            do {
                if (lprof == "gaussian") { do { xv = N01(rng); } while (std::abs(xv) > cut); xv = xi0 + sxi * xv; }
                else xv = xi0 + len * (U01(rng) - 0.5);
            } while (xv < xcut);
            double ppx = spx * N01(rng), ppy = spx * N01(rng);
            const double gam = gamma0 * (1.0 + espread * N01(rng));
            ppx += gam * xp0;
            ppy += gam * yp0;
            const double pz2 = gam * gam - 1.0 - ppx * ppx - ppy * ppy;
            if (pz2 <= 0) throw std::runtime_error(p + ": gamma too small for the transverse momentum spread");
            const double tpx = ppx - gam * xp0, tpy = ppy - gam * yp0;
            for (int l = 0; l < nsym; ++l) {
                const int lr = l % nrot;
                const double sgn = (l >= nrot) ? -1.0 : 1.0;
                const double ph = 2 * PI * lr / nrot, cp = std::cos(ph), sp = std::sin(ph);
                const double qv[7] = {cp * xx - sp * yy + cx0 + cxs * (xv - xi0), sgn * (sp * xx + cp * yy) + cy0 + cys * (xv - xi0),
                                      cp * tpx - sp * tpy + gam * xp0, sgn * (sp * tpx + cp * tpy) + gam * yp0,
                                      std::sqrt(pz2), xv, pwl.empty() ? wmac : wmac * profile(xi_to_user(xv))};
                // Beam::accumulate (src/Beam.cpp:580-588)
                if (qv[6] != 0) {
                    const double gg = std::sqrt(1 + qv[2] * qv[2] + qv[3] * qv[3] + qv[4] * qv[4]);
                    aw += qv[6];
                    awg += qv[6] * gg;
                    awxi += qv[6] * qv[5];
                }
                local.insert(local.end(), qv, qv + 7);
            }
        }
        set_particles(local);
        gref = 0; xiref = 0;
        if (aw > 0) { gref = awg / aw; xiref = awxi / aw; }
        this->pwl = pwl;
        pwl_u0 = pu0; pwl_u1 = pu1; pwl_wmac = wmac;
        xi_init = xi;
    }
    void set_particles(const std::vector<double>& p) {
        Np = static_cast<int>(p.size() / 7);
        x.resize(Np); y.resize(Np); px.resize(Np); py.resize(Np); pz.resize(Np); xi.resize(Np); w.resize(Np);
        for (int i = 0; i < Np; ++i) {
            const double* r = &p[7 * static_cast<size_t>(i)];
            x[i] = r[0]; y[i] = r[1]; px[i] = r[2]; py[i] = r[3]; pz[i] = r[4]; xi[i] = r[5]; w[i] = r[6];
        }
    }
    // src/Beam.cpp:900-905
    template <bool PAR> void push(const Grid& g, const BeamGrid& bg, const Real* fld, int nc, Real dt_kick, Real dt_drift) {
        if (rigid) return;
        const Real qm = q / m;
        if (!started) {
            beam_advance<PAR>(g, bg, m1, qm, mode, 0, Np, x.data(), y.data(), px.data(), py.data(), pz.data(), xi.data(),
                              w.data(), fld, nc, -0.5 * dt_kick, -0.5 * dt_kick, false);
            started = true;
        }
        beam_advance<PAR>(g, bg, m1, qm, mode, 0, Np, x.data(), y.data(), px.data(), py.data(), pz.data(), xi.data(),
                          w.data(), fld, nc, dt_kick, dt_drift, true);
    }
    // src/Beam.cpp:1133-1155
    std::array<double, 17> sums() const {
        std::array<double, 17> S{};
        for (int i = 0; i < Np; ++i) {
            const double ww = w[i];
            if (ww == 0) continue;
            S[16] += 1;
            const double gg = std::sqrt(1 + px[i] * px[i] + py[i] * py[i] + pz[i] * pz[i]);
            const double X = x[i], Y = y[i], PX = px[i], PY = py[i];
            const double G = gg - gref, XI = xi[i] - xiref;
            const double v[15] = {G, G * G, XI, XI * XI, X * X + Y * Y, X, Y, X * X, Y * Y, PX, PY, PX * PX, PY * PY, X * PX, Y * PY};
            S[0] += ww;
            for (int c = 0; c < 15; ++c) S[c + 1] += ww * v[c];
        }
        return S;
    }
    // The same 17 sums in the fixed tree order of the product's on-device reduction (qz_beam_sums_device):
    // blocks of 4096 particles; in a block, lane t (of 256) accumulates particles t, t + 256, ... in index order;
    // the 256 lane partials are halved pairwise (lane t adds lane t + h, h = 128, 64, ..., 1); the block
    // partials are added in block order.  Not in the reference (its Beam::sums is the index-order loop above):
    // it is the summation order the product's beam log and differentiable objective use.
    std::array<double, 17> sums_tree() const {
        constexpr int LANES = 256, CHUNK = 16 * LANES;
        std::array<double, 17> out{};
        for (long lo = 0; lo < Np; lo += CHUNK) {
            const long hi = std::min<long>(Np, lo + CHUNK);
            std::vector<std::array<double, 17>> acc(LANES);
            for (auto& a : acc) a.fill(0.0);
            for (int t = 0; t < LANES; ++t) {
                auto& A = acc[t];
                for (long i = lo + t; i < hi; i += LANES) {
                    const double ww = w[i];
                    if (ww == 0) continue;
                    A[16] += 1;
                    const double PX = px[i], PY = py[i], PZ = pz[i];
                    const double gg = std::sqrt(1 + PX * PX + PY * PY + PZ * PZ);
                    const double X = x[i], Y = y[i];
                    const double G = gg - gref, XI = xi[i] - xiref;
                    const double v[15] = {G, G * G, XI, XI * XI, X * X + Y * Y, X, Y, X * X, Y * Y, PX, PY, PX * PX,
                                          PY * PY, X * PX, Y * PY};
                    A[0] += ww;
                    for (int c = 0; c < 15; ++c) A[c + 1] += ww * v[c];
                }
            }
            for (int h = LANES / 2; h > 0; h >>= 1)
                for (int t = 0; t < h; ++t)
                    for (int c = 0; c < 17; ++c) acc[t][c] += acc[t + h][c];
            for (int c = 0; c < 17; ++c) out[c] += acc[0][c];
        }
        return out;
    }
    // src/Beam.cpp:1157-1192
    BeamDiag from_sums(const std::array<double, 17>& S, Real t) const {
        BeamDiag d;
        d.t = t;
        d.alive = static_cast<long>(S[16]);
        const double sw = S[0];
        if (sw > 0) {
            auto mo = [&](int c) { return S[c + 1] / sw; };
            d.npart = sw;
            d.charge = q * sw;
            d.gamma_mean = gref + mo(0);
            d.gamma_rms = std::sqrt(std::max(0.0, mo(1) - mo(0) * mo(0)));
            d.xi_mean = xi_to_user(xiref + mo(2));
            d.xi_rms = std::sqrt(std::max(0.0, mo(3) - mo(2) * mo(2)));
            d.r_rms = std::sqrt(mo(4));
            d.x_mean = mo(5);
            d.y_mean = mo(6);
            double xx = mo(7) - mo(5) * mo(5), yy = mo(8) - mo(6) * mo(6);
            double pxx = mo(11) - mo(9) * mo(9), pyy = mo(12) - mo(10) * mo(10);
            double xpx = mo(13) - mo(5) * mo(9), ypy = mo(14) - mo(6) * mo(10);
            if (!m1) {
                d.x_mean = d.y_mean = 0;
                xx = yy = 0.5 * mo(4);
                pxx = pyy = 0.5 * (mo(11) + mo(12));
                xpx = ypy = 0.5 * (mo(13) + mo(14));
            }
            d.sigma_x = std::sqrt(std::max(0.0, xx));
            d.sigma_y = std::sqrt(std::max(0.0, yy));
            d.emit_nx = std::sqrt(std::max(0.0, xx * pxx - xpx * xpx));
            d.emit_ny = std::sqrt(std::max(0.0, yy * pyy - ypy * ypy));
        }
        return d;
    }
};

// ============================================================================ simulation (src/Simulation.*)
struct Sim {
    Config cfg;
    bool par = false;
    bool m1 = false;
    int picard = 1;
    Grid grid;
    Solver solver;
    BeamGrid box;
    Real dt = 0, t = 0;
    int nsteps = 0;
    int step = 0;              // next step n
    bool done = false;
    unsigned long long seed = 1;
    bool neutralize = true;
    bool have_prev = false;
    double dt_prev = 0;
    std::vector<Species> species;
    std::vector<Beam> beams;
    SliceSources src;
    SliceFields fld_;
    MScalar bg_rhot, bg_rho;
    MVector gz;
    int M = 0, K = 0, NCF = 0, NCB = 0, NCD = 0;
    std::vector<Real> fld, bsrc, diag;
    bool record_chi = false;
    std::vector<Real> chi_rec;   // (K, M): chi_0 of every slice of the last sweep, when record_chi
    long lost_last = 0;
    double t_sweep = 0, t_beam = 0;
    double particle_slices = 0;   // live mobile particles summed over the slices of the last sweep

    void init(const std::string& deck, const std::string& ov, bool parallel) {
        par = parallel;
        cfg.parse(deck);
        cfg.overrides(ov);
        // src/Simulation.cpp:34-318 (the keys of the path; one rank)
        if (!cfg.has("xi.convention")) throw std::runtime_error("xi.convention is missing");
        if (cfg.get_string("xi.convention") != "z-t") throw std::runtime_error("xi.convention: only z-t");
        const int modes = cfg.get_int("modes", 0);
        if (modes != 0 && modes != 1) throw std::runtime_error("modes must be 0 or 1");
        m1 = (modes == 1);
        picard = cfg.get_int("solver.picard", m1 ? 3 : 1);
        if (cfg.get_double("plasma.smooth_length", 0.0) != 0.0) throw std::runtime_error("oracle: plasma.smooth_length");
        if (cfg.get_double("pusher.max_cells_per_step", 0.0) != 0.0) throw std::runtime_error("oracle: sub-slicing");
        grid.finalize(grid_nodes(cfg));
        solver.init(grid);
        const Real xu_max = cfg.get_double("xi.max", 0.0);
        const Real xu_min = cfg.get_double("xi.min");
        box.xi_min = xi_from_user(xu_max);
        box.dxi = cfg.get_double("xi.step");
        if (!(xu_max > xu_min) || !(box.dxi > 0)) throw std::runtime_error("xi: need xi.min < xi.max and xi.step > 0");
        box.nxi = static_cast<int>(std::lround((xu_max - xu_min) / box.dxi)) + 1;
        dt = cfg.get_double("time.dt", 0.0);
        nsteps = cfg.get_int("time.steps", 0);
        t = cfg.get_double("time.start", 0.0);
        if (dt <= 0 && nsteps > 0) throw std::runtime_error("time.dt must be > 0");
        seed = cfg.get_int("plasma.seed", 1);
        box.k0 = 0;
        box.nloc = box.nxi;
        const std::string shape = cfg.get_string("beams.xi_shape", "linear");
        if (shape == "ngp") box.ngp = true;
        else if (shape == "linear") box.ngp = false;
        else throw std::runtime_error("beams.xi_shape must be linear or ngp");
        auto names = cfg.get_list("plasma.species");
        if (names.empty()) names.push_back("electrons");
        bool has_positive = false;
        species.resize(names.size());
        for (size_t i = 0; i < names.size(); ++i) {
            species[i].init(cfg, names[i], grid, m1);
            if (species[i].q > 0) has_positive = true;
        }
        neutralize = cfg.get_bool("plasma.neutralize", !has_positive);
        const auto bnames = cfg.get_list("beams");
        beams.resize(bnames.size());
        for (size_t i = 0; i < bnames.size(); ++i) beams[i].init(cfg, bnames[i], m1);
        M = grid.N + 1;
        K = box.nloc;
        NCF = m1 ? F_NC1 : F_NC0;
        NCB = m1 ? B_NC1 : B_NC0;
        NCD = m1 ? D_NC1 : D_NC0;
        src.init(M);
        fld_.init(M);
        bg_rhot.init(M);
        bg_rho.init(M);
        gz.init(M);
        fld.assign(static_cast<size_t>(K) * M * NCF, 0);
        bsrc.assign(static_cast<size_t>(K) * M * NCB, 0);
        diag.assign(static_cast<size_t>(K) * M * NCD, 0);
        step = cfg.get_int("time.first_step", 0);
    }

    // src/Simulation.cpp:394-406
    void zero_sources() { src.init(M); }

    // src/Simulation.cpp:408-457 (ds = 0, one rank)
    void combine_sources(int kl) {
        const int KL = box.nloc;
        const int k = kl;
        SliceSources& s = src;
        const Real idxi = Real(1) / box.dxi;
        const bool ngp = box.ngp;
        auto b = [&](int kk, int j, int comp) { return bsrc[(static_cast<size_t>(kk) * M + j) * NCB + comp]; };
        for (int j = 0; j < M; ++j) {
            auto dxi_b = [&](int comp) {
                if (ngp) {
                    const Real prev = (k > 0) ? b(k - 1, j, comp) : Real(0);
                    return (b(k, j, comp) - prev) * idxi;
                }
                if (KL == 1) return Real(0);
                if (k == 0) return (b(1, j, comp) - b(0, j, comp)) * idxi;
                if (k == KL - 1) return (b(KL - 1, j, comp) - b(KL - 2, j, comp)) * idxi;
                return Real(0.5) * (b(k + 1, j, comp) - b(k - 1, j, comp)) * idxi;
            };
            auto bv = [&](int comp) { return b(k, j, comp); };
            s.rhot.a0[j] += bg_rhot.a0[j] + bv(B_RT0);
            s.drho.a0[j] += dxi_b(B_RT0);
            s.jz.a0[j] += bv(B_JZ0);
            s.jp.r1[j] += bv(B_JP1R);
            s.jp.i1[j] += bv(B_JP1I);
            s.rho.a0[j] += bg_rho.a0[j] + bv(B_RHO0);
            if (bg_rho.a0[j] > Real(0)) s.ni.a0[j] += bg_rho.a0[j];
            if (m1) {
                s.rhot.ar[j] += bg_rhot.ar[j] + bv(B_RT1R);
                s.rhot.ai[j] += bg_rhot.ai[j] + bv(B_RT1I);
                s.drho.ar[j] += dxi_b(B_RT1R);
                s.drho.ai[j] += dxi_b(B_RT1I);
                s.jz.ar[j] += bv(B_JZ1R);
                s.jz.ai[j] += bv(B_JZ1I);
                s.jp.r0[j] += bv(B_JP0R);
                s.jp.i0[j] += bv(B_JP0I);
                s.jp.r2[j] += bv(B_JP2R);
                s.jp.i2[j] += bv(B_JP2I);
                s.rho.ar[j] += bg_rho.ar[j] + bv(B_RHO1R);
                s.rho.ai[j] += bg_rho.ai[j] + bv(B_RHO1I);
            }
        }
    }

    // src/Simulation.cpp:477-498
    void compute_wplus() {
        SliceFields& f = fld_;
        solver.gradient(f.psi.a0.data(), f.t1.data());
        if (m1) {
            solver.gradient_odd(f.psi.ar.data(), f.t2.data());
            solver.over_r(f.psi.ar.data(), f.t3.data());
            solver.gradient_odd(f.psi.ai.data(), f.t4.data());
            solver.over_r(f.psi.ai.data(), f.rhs.data());
        }
        for (int j = 0; j < M; ++j) {
            f.wp.r1[j] = -f.t1[j];
            f.wp.i1[j] = Real(0);
            if (m1) {
                f.wp.r2[j] = -(f.t2[j] - f.t3[j]);
                f.wp.i2[j] = -(f.t4[j] - f.rhs[j]);
                f.wp.r0[j] = -(f.t2[j] + f.t3[j]);
                f.wp.i0[j] = (f.t4[j] + f.rhs[j]);
            }
        }
    }

    // src/Simulation.cpp:501-532
    void compute_bz() {
        SliceFields& f = fld_;
        SliceSources& s = src;
        MVector& g = gz;
        solver.gradient_odd(s.jp.i1.data(), f.t1.data());
        solver.over_r(s.jp.i1.data(), f.t2.data());
        if (m1) {
            solver.gradient_odd(s.jp.r2.data(), g.r0.data());
            solver.over_r(s.jp.r2.data(), g.i0.data());
            solver.gradient_odd(s.jp.i2.data(), g.r1.data());
            solver.over_r(s.jp.i2.data(), g.i1.data());
            solver.gradient(s.jp.r0.data(), g.r2.data());
            solver.gradient(s.jp.i0.data(), g.i2.data());
        }
        for (int j = 0; j < M; ++j) {
            f.t3[j] = -(f.t1[j] + f.t2[j]);
            if (m1) {
                const Real D1r = g.r0[j] + Real(2) * g.i0[j], D1i = g.r1[j] + Real(2) * g.i1[j];
                const Real Dmr = g.r2[j], Dmi = g.i2[j];
                f.t4[j] = -Real(0.5) * (D1i + Dmi);
                f.rhs[j] = Real(0.5) * (D1r - Dmr);
            }
        }
        solver.solve(L0, f.t3.data(), f.bz.a0.data());
        if (m1) {
            solver.solve(L1D, f.t4.data(), f.bz.ar.data());
            solver.solve(L1D, f.rhs.data(), f.bz.ai.data());
        }
    }

    // src/Simulation.cpp:535-591
    void solve_bplus() {
        SliceFields& f = fld_;
        SliceSources& s = src;
        MVector& g = gz;
        solver.gradient(s.jz.a0.data(), f.t1.data());
        if (m1) {
            solver.gradient_odd(s.jz.ar.data(), f.t2.data());
            solver.over_r(s.jz.ar.data(), f.t3.data());
            solver.gradient_odd(s.jz.ai.data(), f.t4.data());
            solver.over_r(s.jz.ai.data(), f.rhs.data());
        }
        for (int j = 0; j < M; ++j) {
            g.r1[j] = f.t1[j] + s.S.r1[j];
            g.i1[j] = s.S.i1[j];
            if (m1) {
                g.r2[j] = (f.t2[j] - f.t3[j]) + s.S.r2[j];
                g.i2[j] = (f.t4[j] - f.rhs[j]) + s.S.i2[j];
                g.r0[j] = (f.t2[j] + f.t3[j]) + s.S.r0[j];
                g.i0[j] = -(f.t4[j] + f.rhs[j]) + s.S.i0[j];
            }
        }
        MVector& rb = f.bprev;
        const int niter = m1 ? picard : 1;
        for (int it = 0; it < niter; ++it) {
            for (int j = 0; j < M; ++j) {
                Real r1 = -g.i1[j], i1 = g.r1[j];
                if (m1) {
                    Real r0 = -g.i0[j], i0 = g.r0[j], r2 = -g.i2[j], i2 = g.r2[j];
                    const Real cr = s.chi.ar[j], ci = s.chi.ai[j];
                    const Real b0r = f.bp.r0[j], b0i = f.bp.i0[j], b1r = f.bp.r1[j], b1i = f.bp.i1[j];
                    const Real b2r = f.bp.r2[j], b2i = f.bp.i2[j];
                    r0 += cr * b1r + ci * b1i; i0 += cr * b1i - ci * b1r;
                    r1 += cr * b0r - ci * b0i + cr * b2r + ci * b2i;
                    i1 += cr * b0i + ci * b0r + cr * b2i - ci * b2r;
                    r2 += cr * b1r - ci * b1i; i2 += cr * b1i + ci * b1r;
                    rb.r0[j] = r0; rb.i0[j] = i0; rb.r2[j] = r2; rb.i2[j] = i2;
                }
                rb.r1[j] = r1; rb.i1[j] = i1;
            }
            solver.solve(L1D, rb.r1.data(), f.bp.r1.data(), s.chi.a0.data());
            solver.solve(L1F, rb.i1.data(), f.bp.i1.data(), s.chi.a0.data(), s.jz.a0.data());
            if (m1) {
                solver.solve(L0, rb.r0.data(), f.bp.r0.data(), s.chi.a0.data());
                solver.solve(L0, rb.i0.data(), f.bp.i0.data(), s.chi.a0.data());
                solver.solve(L2D, rb.r2.data(), f.bp.r2.data(), s.chi.a0.data());
                solver.solve(L2D, rb.i2.data(), f.bp.i2.data(), s.chi.a0.data());
            }
        }
    }

    // src/Simulation.cpp:797-819 (no smoothing filter)
    template <bool PAR> void solve_slice_fields() {
        solver.solve(L0, src.rhot.a0.data(), fld_.psi.a0.data(), nullptr, nullptr, -1.0);
        solver.solve(L0, src.drho.a0.data(), fld_.ez.a0.data(), nullptr, nullptr, -1.0);
        if (m1) {
            solver.solve(L1D, src.rhot.ar.data(), fld_.psi.ar.data(), nullptr, nullptr, -1.0);
            solver.solve(L1D, src.rhot.ai.data(), fld_.psi.ai.data(), nullptr, nullptr, -1.0);
            solver.solve(L1D, src.drho.ar.data(), fld_.ez.ar.data(), nullptr, nullptr, -1.0);
            solver.solve(L1D, src.drho.ai.data(), fld_.ez.ai.data(), nullptr, nullptr, -1.0);
        }
        compute_wplus();
        compute_bz();
        for (auto& sp : species)
            if (sp.mobile) sp.template deposit_S<PAR>(fld_, src.S);
        solve_bplus();
    }

    // src/Simulation.cpp:594-633
    void store_slice(int kl) {
        const int k = kl;
        SliceFields& f = fld_;
        SliceSources& s = src;
        auto F = [&](int j, int comp) -> Real& { return fld[(static_cast<size_t>(k) * M + j) * NCF + comp]; };
        auto D = [&](int j, int comp) -> Real& { return diag[(static_cast<size_t>(k) * M + j) * NCD + comp]; };
        for (int j = 0; j < M; ++j) {
            F(j, F_EP1R) = f.wp.r1[j] + f.bp.i1[j];
            F(j, F_EP1I) = f.wp.i1[j] - f.bp.r1[j];
            F(j, F_BP1R) = f.bp.r1[j];
            F(j, F_BP1I) = f.bp.i1[j];
            F(j, F_EZ0) = f.ez.a0[j];
            F(j, F_BZ0) = f.bz.a0[j];
            D(j, D_PSI0) = f.psi.a0[j];
            D(j, D_NE0) = s.ne.a0[j];
            D(j, D_NI0) = s.ni.a0[j];
            if (m1) {
                F(j, F_EP0R) = f.wp.r0[j] + f.bp.i0[j];
                F(j, F_EP0I) = f.wp.i0[j] - f.bp.r0[j];
                F(j, F_EP2R) = f.wp.r2[j] + f.bp.i2[j];
                F(j, F_EP2I) = f.wp.i2[j] - f.bp.r2[j];
                F(j, F_BP0R) = f.bp.r0[j];
                F(j, F_BP0I) = f.bp.i0[j];
                F(j, F_BP2R) = f.bp.r2[j];
                F(j, F_BP2I) = f.bp.i2[j];
                F(j, F_EZ1R) = f.ez.ar[j];
                F(j, F_EZ1I) = f.ez.ai[j];
                F(j, F_BZ1R) = f.bz.ar[j];
                F(j, F_BZ1I) = f.bz.ai[j];
                D(j, D_PSI1R) = f.psi.ar[j];
                D(j, D_PSI1I) = f.psi.ai[j];
                D(j, D_NE1R) = s.ne.ar[j];
                D(j, D_NE1I) = s.ne.ai[j];
                D(j, D_NI1R) = s.ni.ar[j];
                D(j, D_NI1I) = s.ni.ai[j];
            }
        }
    }

    // the beam part of step_fields (src/Simulation.cpp:642-643)
    template <bool PAR> void deposit_beams() {
        std::fill(bsrc.begin(), bsrc.end(), 0.0);
        for (auto& b : beams) {
            if (b.analytic) beam_deposit_analytic(grid, box, m1, b.q, b.an, bsrc.data(), NCB);
            else beam_deposit<PAR>(grid, box, m1, b.q, b.Np, b.x.data(), b.y.data(), b.px.data(), b.py.data(),
                                   b.pz.data(), b.xi.data(), b.w.data(), bsrc.data(), NCB);
        }
    }

    // plasma load, background and the B+ warm start of step_fields (src/Simulation.cpp:653-708)
    template <bool PAR> void load_plasma() {
        bg_rhot.init(M);
        bg_rho.init(M);
        const Real z_front = t - box.xi_min;
        unsigned long long sd = seed;
        for (auto& s : species) s.load(z_front, sd++);
        for (auto& s : species)
            if (!s.mobile) s.template deposit_rho<PAR>(bg_rhot, bg_rho);
        if (neutralize) {
            MScalar& a = src.rhot;
            MScalar& b = src.rho;
            a.init(M);
            b.init(M);
            for (auto& s : species)
                if (s.mobile) s.template deposit_rho<PAR>(a, b);
            for (int j = 0; j < M; ++j) {
                bg_rhot.a0[j] -= a.a0[j]; bg_rhot.ar[j] -= a.ar[j]; bg_rhot.ai[j] -= a.ai[j];
                bg_rho.a0[j] -= b.a0[j]; bg_rho.ar[j] -= b.ar[j]; bg_rho.ai[j] -= b.ai[j];
            }
        }
        fld_.bp.init(M);
    }

    // the slice loop of step_fields (src/Simulation.cpp:727-783; no sub-slicing, one rank)
    template <bool PAR> void sweep_slices() {
        const int Kg = box.nxi, k0 = box.k0, KL = box.nloc;
        const Real dxi = box.dxi;
        particle_slices = 0;
        for (int kl = 0; kl < KL; ++kl) {
            const int k = k0 + kl;
            for (auto& sp : species)
                if (sp.mobile)
                    for (int i = 0; i < sp.Np; ++i) particle_slices += sp.w[i] != Real(0) ? 1.0 : 0.0;
            zero_sources();
            for (auto& sp : species)
                if (sp.mobile) sp.template deposit<PAR>(src);
            combine_sources(kl);
            if (record_chi)   // chi_0 of the slice: what advance_slice of the laser takes (src/Simulation.cpp:735)
                std::copy(src.chi.a0.begin(), src.chi.a0.end(), chi_rec.begin() + static_cast<size_t>(kl) * M);
            solve_slice_fields<PAR>();
            store_slice(kl);
            if (k < Kg - 1)
                for (auto& sp : species)
                    if (sp.mobile) sp.template push<PAR>(fld_, dxi);
        }
    }

    // src/Simulation.cpp:636-795
    // test hooks of the adjoint checks: added to the beam sources after the deposit and to the fields after
    // the sweep (empty: no change)
    std::vector<Real> bsrc_delta, fld_delta;
    std::vector<double> adj_lf, adj_lb, adj_lw;   // the adjoints of the last or_sim_spread_grad
    void add_delta(std::vector<Real>& a, const std::vector<Real>& d) const {
        if (!d.empty())
            for (size_t i = 0; i < a.size(); ++i) a[i] += d[i];
    }

    template <bool PAR> void step_fields() {
        deposit_beams<PAR>();
        add_delta(bsrc, bsrc_delta);
        load_plasma<PAR>();
        sweep_slices<PAR>();
        add_delta(fld, fld_delta);
    }

    // one iteration of Simulation::run (src/Simulation.cpp:1376-1459) without output writing
    template <bool PAR> void do_step() {
        const int n = step;
        const bool last = (n >= nsteps);
        double t0 = now();
        step_fields<PAR>();
        t_sweep = now() - t0;
        lost_last = 0;
        for (auto& s : species)
            if (s.mobile) { lost_last += s.lost; s.lost = 0; }
        t_beam = 0;
        if (!last) {
            if (!have_prev) { dt_prev = dt; have_prev = true; }
            const Real dkick = Real(0.5) * (dt_prev + dt), ddrift = dt;
            t0 = now();
            for (auto& b : beams) b.template push<PAR>(grid, box, fld.data(), NCF, dkick, ddrift);
            t_beam = now() - t0;
            t += dt;
            dt_prev = dt;
        } else {
            done = true;
        }
        ++step;
    }
    static double now() {
#ifdef _OPENMP
        return omp_get_wtime();
#else
        return 0.0;
#endif
    }
};

// fields_*.bin version 3 (src/Simulation.cpp:873-993, main group with all components)
static void write_fields(const Sim& s, const std::string& fn) {
    const int M = s.M, K = s.K;
    auto F = [&](int k, int j, int c) { return s.fld[(static_cast<size_t>(k) * M + j) * s.NCF + c]; };
    auto D = [&](int k, int j, int c) { return s.diag[(static_cast<size_t>(k) * M + j) * s.NCD + c]; };
    auto B = [&](int k, int j, int c) { return s.bsrc[(static_cast<size_t>(k) * M + j) * s.NCB + c]; };
    std::vector<std::pair<std::string, std::function<double(int, int)>>> comps;
    auto addc = [&](const std::string& n, std::function<double(int, int)> f) { comps.emplace_back(n, f); };
    addc("psi", [&](int k, int j) { return D(k, j, D_PSI0); });
    addc("ez", [&](int k, int j) { return F(k, j, F_EZ0); });
    addc("er", [&](int k, int j) { return F(k, j, F_EP1R); });
    addc("eth", [&](int k, int j) { return F(k, j, F_EP1I); });
    addc("br", [&](int k, int j) { return F(k, j, F_BP1R); });
    addc("bth", [&](int k, int j) { return F(k, j, F_BP1I); });
    addc("bz", [&](int k, int j) { return F(k, j, F_BZ0); });
    addc("ne", [&](int k, int j) { return D(k, j, D_NE0); });
    addc("ni", [&](int k, int j) { return D(k, j, D_NI0); });
    addc("rhob", [&](int k, int j) { return B(k, j, B_RHO0); });
    if (s.m1) {
        auto scalD = [&](const std::string& n, int cr, int ci) {
            addc(n + "_c", [&, cr](int k, int j) { return 2.0 * D(k, j, cr); });
            addc(n + "_s", [&, ci](int k, int j) { return -2.0 * D(k, j, ci); });
        };
        auto scalF = [&](const std::string& n, int cr, int ci) {
            addc(n + "_c", [&, cr](int k, int j) { return 2.0 * F(k, j, cr); });
            addc(n + "_s", [&, ci](int k, int j) { return -2.0 * F(k, j, ci); });
        };
        scalD("psi", D_PSI1R, D_PSI1I);
        scalF("ez", F_EZ1R, F_EZ1I);
        scalF("bz", F_BZ1R, F_BZ1I);
        scalD("ne", D_NE1R, D_NE1I);
        scalD("ni", D_NI1R, D_NI1I);
        addc("rhob_c", [&](int k, int j) { return 2.0 * B(k, j, B_RHO1R); });
        addc("rhob_s", [&](int k, int j) { return -2.0 * B(k, j, B_RHO1I); });
        auto vec = [&](const std::string& nr, const std::string& nt, int c0r, int c0i, int c2r, int c2i) {
            addc(nr + "_c", [&, c0r, c2r](int k, int j) { return F(k, j, c0r) + F(k, j, c2r); });
            addc(nr + "_s", [&, c0i, c2i](int k, int j) { return F(k, j, c0i) - F(k, j, c2i); });
            addc(nt + "_c", [&, c0i, c2i](int k, int j) { return F(k, j, c0i) + F(k, j, c2i); });
            addc(nt + "_s", [&, c0r, c2r](int k, int j) { return F(k, j, c2r) - F(k, j, c0r); });
        };
        vec("er", "eth", F_EP0R, F_EP0I, F_EP2R, F_EP2I);
        vec("br", "bth", F_BP0R, F_BP0I, F_BP2R, F_BP2I);
    }
    std::FILE* fp = std::fopen(fn.c_str(), "wb");
    if (!fp) throw std::runtime_error("cannot write " + fn);
    const int32_t hdr[4] = {3, M, K, s.m1 ? 1 : 0};
    std::fwrite(hdr, 4, 4, fp);
    std::fwrite(&s.t, 8, 1, fp);
    std::fwrite(s.grid.r.data(), 8, M, fp);
    for (int k = 0; k < K; ++k) {
        const double xx = xi_to_user(s.box.xi_min + (K - 1 - k) * 1 * s.box.dxi);
        std::fwrite(&xx, 8, 1, fp);
    }
    const int32_t nc = static_cast<int32_t>(comps.size());
    std::fwrite(&nc, 4, 1, fp);
    std::vector<double> v(static_cast<size_t>(M));
    for (auto& c : comps) {
        char name[16];
        std::memset(name, 0, sizeof(name));
        std::strncpy(name, c.first.c_str(), 15);
        std::fwrite(name, 1, 16, fp);
        for (int row = 0; row < K; ++row) {
            const int k = K - 1 - row;
            for (int j = 0; j < M; ++j) v[j] = c.second(k, j);
            std::fwrite(v.data(), 8, M, fp);
        }
    }
    std::fclose(fp);
}

// src/Simulation.cpp:995-1029 (one rank): fixed-width text, a 256-byte header of two padded lines, then one line
// per slice from the tail to the head: xi, Ez(0), psi(0), ne(0), rho_beam(0), (Er - Bth)/r at the first node
// off the axis, and with m = 1 the axis values of W = E_perp + z x B_perp.
static void write_axis(const Sim& s, const std::string& fn) {
    const int M = s.M, K = s.K;
    auto F = [&](int k, int j, int c) { return s.fld[(static_cast<size_t>(k) * M + j) * s.NCF + c]; };
    auto D = [&](int k, int j, int c) { return s.diag[(static_cast<size_t>(k) * M + j) * s.NCD + c]; };
    auto B = [&](int k, int j, int c) { return s.bsrc[(static_cast<size_t>(k) * M + j) * s.NCB + c]; };
    const double r1 = s.grid.r[1];
    const int ncol = s.m1 ? 8 : 6;
    std::string h = "# t = " + std::to_string(s.t);
    h.resize(127, ' ');
    h += '\n';
    std::string h2 = "# xi  Ez(0)  psi(0)  ne(0)  rho_beam(0)  (Er-Bth)/r|r1";
    if (s.m1) h2 += "  Wx(0)  Wy(0)  [W = E_perp + z x B_perp on the axis, from m=1]";
    h2.resize(127, ' ');
    h2 += '\n';
    h += h2;
    std::string buf;
    char cell[32];
    for (int k = K - 1; k >= 0; --k) {
        double v[8] = {xi_to_user(s.box.xi_min + k * s.box.dxi), F(k, 0, F_EZ0), D(k, 0, D_PSI0), D(k, 0, D_NE0),
                       B(k, 0, B_RHO0), (F(k, 1, F_EP1R) - F(k, 1, F_BP1I)) / r1, 0, 0};
        if (s.m1) { v[6] = F(k, 0, F_EP0R) - F(k, 0, F_BP0I); v[7] = F(k, 0, F_EP0I) + F(k, 0, F_BP0R); }
        for (int c = 0; c < ncol; ++c) {
            std::snprintf(cell, sizeof(cell), "%+.10e%c", v[c], c == ncol - 1 ? '\n' : ' ');
            buf += cell;
        }
    }
    std::FILE* fp = std::fopen(fn.c_str(), "wb");
    if (!fp) throw std::runtime_error("cannot write " + fn);
    std::fwrite(h.data(), 1, h.size(), fp);
    std::fwrite(buf.data(), 1, buf.size(), fp);
    std::fclose(fp);
}

}  // namespace orc

// ============================================================================ C API (test infrastructure)
using namespace orc;

#define OR_TRY(...)                                    \
    try { __VA_ARGS__; return 0; }                     \
    catch (const std::exception& e) { last_error = e.what(); return -1; }

#include "adjoint_oracle.inc"

extern "C" {

const char* or_last_error() { return last_error.c_str(); }

int or_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

int or_thomas(const double* a, const double* b, const double* c, const double* d, double* x, int M) {
    std::vector<double> cp(M), dp(M);
    thomas(a, b, c, d, x, M, cp.data(), dp.data());
    return 0;
}

int or_make_uniform(double R, double dr, double* out, int cap) {
    try {
        auto v = make_uniform(R, dr);
        if (static_cast<int>(v.size()) > cap) return -static_cast<int>(v.size());
        std::copy(v.begin(), v.end(), out);
        return static_cast<int>(v.size());
    } catch (const std::exception& e) { last_error = e.what(); return 0; }
}

int or_make_regions(double R, const double* ends, const double* hs, int nreg, double max_ratio, double* out, int cap) {
    try {
        std::vector<std::pair<double, double>> reg;
        for (int i = 0; i < nreg; ++i) reg.emplace_back(ends[i], hs[i]);
        auto v = make_regions(R, reg, max_ratio);
        if (static_cast<int>(v.size()) > cap) return -static_cast<int>(v.size());
        std::copy(v.begin(), v.end(), out);
        return static_cast<int>(v.size());
    } catch (const std::exception& e) { last_error = e.what(); return 0; }
}

// grid + field solver on an explicit node list
void* or_solver_create(const double* nodes, int M) {
    try {
        struct GS { Grid g; Solver s; };
        auto* gs = new GS;
        gs->g.finalize(std::vector<Real>(nodes, nodes + M));
        gs->s.init(gs->g);
        return gs;
    } catch (const std::exception& e) { last_error = e.what(); return nullptr; }
}
void or_solver_destroy(void* h) {
    struct GS { Grid g; Solver s; };
    delete static_cast<GS*>(h);
}
static Solver& S_of(void* h) {
    struct GS { Grid g; Solver s; };
    return static_cast<GS*>(h)->s;
}
static Grid& G_of(void* h) {
    struct GS { Grid g; Solver s; };
    return static_cast<GS*>(h)->g;
}
int or_solver_solve(void* h, int op, const double* rhs, double* f, const double* chi, const double* wall, double scale) {
    OR_TRY(S_of(h).solve(op, rhs, f, chi, wall, scale))
}
int or_solver_gradient(void* h, const double* f, double* df, int odd) {
    OR_TRY(if (odd) S_of(h).gradient_odd(f, df); else S_of(h).gradient(f, df))
}
int or_solver_over_r(void* h, const double* f, double* out) { OR_TRY(S_of(h).over_r(f, out)) }
int or_solver_coeffs(void* h, int op, double* a, double* b, double* c) {
    OR_TRY({
        const Op& o = S_of(h).ops[op];
        std::copy(o.a.begin(), o.a.end(), a);
        std::copy(o.b.begin(), o.b.end(), b);
        std::copy(o.c.begin(), o.c.end(), c);
    })
}
double or_solver_flux_coef(void* h) { return S_of(h).flux_coef; }
int or_grid_arrays(void* h, double* V, double* inv_V, double* inv_h) {
    OR_TRY({
        const Grid& g = G_of(h);
        std::copy(g.V.begin(), g.V.end(), V);
        std::copy(g.inv_V.begin(), g.inv_V.end(), inv_V);
        std::copy(g.inv_h.begin(), g.inv_h.end(), inv_h);
    })
}
int or_grid_locate(void* h, int n, const double* xs, int* j, double* t) {
    OR_TRY(for (int i = 0; i < n; ++i) G_of(h).weights(xs[i], j[i], t[i]))
}

// uniform plasma deposit of one species on a grid (tests/test_solvers.cpp:126-143)
int or_uniform_deposit(void* h, int ppc, int ntheta, int mode1, double* rhot0) {
    OR_TRY({
        Config cfg;
        cfg.kv["electrons.ppc"] = std::to_string(ppc);
        cfg.kv["electrons.ntheta"] = std::to_string(ntheta);
        Species sp;
        sp.init(cfg, "electrons", G_of(h), mode1 != 0);
        sp.load(0.0, 1);
        MScalar rt, rh;
        rt.init(G_of(h).N + 1);
        rh.init(G_of(h).N + 1);
        sp.deposit_rho<false>(rt, rh);
        std::copy(rt.a0.begin(), rt.a0.end(), rhot0);
    })
}

// pushers on n particles (u, E, B: n x 3)
int or_push_batch(int mode, int n, const double* u, const double* E, const double* B, const double* eps, double* out) {
    for (int i = 0; i < n; ++i) {
        const V3 uu{u[3 * i], u[3 * i + 1], u[3 * i + 2]};
        const V3 EE{E[3 * i], E[3 * i + 1], E[3 * i + 2]};
        const V3 BB{B[3 * i], B[3 * i + 1], B[3 * i + 2]};
        const V3 r = push_u(mode, uu, EE, BB, eps[i]);
        out[3 * i] = r.x; out[3 * i + 1] = r.y; out[3 * i + 2] = r.z;
    }
    return 0;
}

int or_make_ab(int order, double* c) {
    OR_TRY({ ABCoeffs a = make_ab(order); std::copy(a.c, a.c + 5, c); })
}
int or_make_ab_variable(int order, const double* t, double h, double* c) {
    OR_TRY({ ABCoeffs a = make_ab_variable(order, t, h); std::copy(a.c, a.c + 5, c); })
}

unsigned long long or_mt19937_64(unsigned long long seed, int n) {
    std::mt19937_64 rng(seed);
    unsigned long long v = 0;
    for (int i = 0; i < n; ++i) v = rng();
    return v;
}

// beam kernels on explicit arrays (grid from a solver handle)
int or_beam_deposit(void* h, double xi_min, double dxi, int nxi, int ngp, int m1, double q, int Np,
                    const double* x, const double* y, const double* px, const double* py, const double* pz,
                    const double* xi, const double* w, double* src, int nc) {
    OR_TRY({
        BeamGrid bg; bg.xi_min = xi_min; bg.dxi = dxi; bg.nxi = nxi; bg.k0 = 0; bg.nloc = nxi; bg.ngp = ngp != 0;
        beam_deposit<false>(G_of(h), bg, m1 != 0, q, Np, x, y, px, py, pz, xi, w, src, nc);
    })
}
int or_beam_advance(void* h, double xi_min, double dxi, int nxi, int ngp, int m1, double qm, int mode, int Np,
                    double* x, double* y, double* px, double* py, double* pz, double* xi, double* w,
                    const double* fld, int nc, double dt_kick, double dt, int move) {
    OR_TRY({
        BeamGrid bg; bg.xi_min = xi_min; bg.dxi = dxi; bg.nxi = nxi; bg.k0 = 0; bg.nloc = nxi; bg.ngp = ngp != 0;
        beam_advance<false>(G_of(h), bg, m1 != 0, qm, mode, 0, Np, x, y, px, py, pz, xi, w, fld, nc, dt_kick, dt, move != 0);
    })
}

// ---- whole simulation
void* or_sim_create(const char* deck, const char* overrides, int parallel) {
    try {
        auto* s = new Sim;
        s->init(deck, overrides ? overrides : "", parallel != 0);
        return s;
    } catch (const std::exception& e) { last_error = e.what(); return nullptr; }
}
void or_sim_destroy(void* h) { delete static_cast<Sim*>(h); }

// info: M K m1 NCF NCB NCD nspecies nbeams nsteps step
int or_sim_info(void* h, int* info) {
    Sim& s = *static_cast<Sim*>(h);
    info[0] = s.M; info[1] = s.K; info[2] = s.m1; info[3] = s.NCF; info[4] = s.NCB; info[5] = s.NCD;
    info[6] = static_cast<int>(s.species.size()); info[7] = static_cast<int>(s.beams.size());
    info[8] = s.nsteps; info[9] = s.step;
    return 0;
}
int or_sim_step(void* h) {
    Sim& s = *static_cast<Sim*>(h);
    OR_TRY(if (s.par) s.do_step<true>(); else s.do_step<false>())
}
// the sweep only (beam deposit, load, slice loop), for timing
int or_sim_step_fields(void* h) {
    Sim& s = *static_cast<Sim*>(h);
    OR_TRY({
        const double t0 = Sim::now();
        if (s.par) s.step_fields<true>(); else s.step_fields<false>();
        s.t_sweep = Sim::now() - t0;
    })
}
// the first nk slices of a sweep (after deposit_beams/load_plasma), for a bounded CPU-baseline sample
int or_sim_sweep_partial(void* h, int nk) {
    Sim& s = *static_cast<Sim*>(h);
    OR_TRY({
        s.deposit_beams<false>();
        s.load_plasma<false>();
        const int save = s.box.nloc;
        const int save_nxi = s.box.nxi;
        s.box.nloc = std::min(nk, save);
        const double t0 = Sim::now();
        if (s.par) s.sweep_slices<true>(); else s.sweep_slices<false>();
        s.t_sweep = Sim::now() - t0;
        s.box.nloc = save;
        s.box.nxi = save_nxi;
    })
}
int or_sim_timing(void* h, double* out) {
    Sim& s = *static_cast<Sim*>(h);
    out[0] = s.t_sweep; out[1] = s.t_beam; out[2] = s.particle_slices;
    return 0;
}
long or_sim_lost(void* h) { return static_cast<Sim*>(h)->lost_last; }
double or_sim_time(void* h) { return static_cast<Sim*>(h)->t; }

// what: fld diag bsrc nodes; beam arrays "beam<b>.<x|y|px|py|pz|xi|w>"; species "sp<s>.<x|y|px|py|dl|w>"
long or_sim_get(void* h, const char* what, double* out, long cap) {
    Sim& s = *static_cast<Sim*>(h);
    const std::vector<double>* v = nullptr;
    std::string w(what);
    if (w == "fld") v = &s.fld;
    else if (w == "chi") v = &s.chi_rec;
    else if (w == "diag") v = &s.diag;
    else if (w == "bsrc") v = &s.bsrc;
    else if (w == "nodes") v = &s.grid.r;
    else if (w.rfind("beam", 0) == 0) {
        const int b = std::atoi(w.c_str() + 4);
        const std::string c = w.substr(w.find('.') + 1);
        if (b < 0 || b >= static_cast<int>(s.beams.size())) return -1;
        Beam& B = s.beams[b];
        v = c == "x" ? &B.x : c == "y" ? &B.y : c == "px" ? &B.px : c == "py" ? &B.py : c == "pz" ? &B.pz : c == "xi" ? &B.xi : c == "w" ? &B.w : nullptr;
    } else if (w.rfind("sp", 0) == 0) {
        const int b = std::atoi(w.c_str() + 2);
        const std::string c = w.substr(w.find('.') + 1);
        if (b < 0 || b >= static_cast<int>(s.species.size())) return -1;
        Species& S = s.species[b];
        v = c == "x" ? &S.x : c == "y" ? &S.y : c == "px" ? &S.px : c == "py" ? &S.py : c == "dl" ? &S.dl : c == "w" ? &S.w : nullptr;
        if (v) {
            const long n = S.Np;
            if (out && cap >= n) std::copy(v->begin(), v->begin() + n, out);
            return n;
        }
    }
    if (!v) return -1;
    const long n = static_cast<long>(v->size());
    if (out && cap >= n) std::copy(v->begin(), v->end(), out);
    return n;
}

int or_sim_beam_sums(void* h, int b, double* out17, double* refs) {
    Sim& s = *static_cast<Sim*>(h);
    if (b < 0 || b >= static_cast<int>(s.beams.size())) return -1;
    const auto S = s.beams[b].sums();
    std::copy(S.begin(), S.end(), out17);
    if (refs) { refs[0] = s.beams[b].gref; refs[1] = s.beams[b].xiref; }
    return 0;
}
// BeamDiag: t npart charge gamma_mean gamma_rms xi_mean xi_rms r_rms emit_nx x_mean y_mean emit_ny sigma_x sigma_y alive
// tree = 0: from the index-order sums (the reference's); tree = 1: from the on-device reduction's order
int or_sim_beam_diag(void* h, int b, int tree, double* out15) {
    Sim& s = *static_cast<Sim*>(h);
    if (b < 0 || b >= static_cast<int>(s.beams.size())) return -1;
    const BeamDiag d = s.beams[b].from_sums(tree ? s.beams[b].sums_tree() : s.beams[b].sums(), s.t);
    const double v[15] = {d.t, d.npart, d.charge, d.gamma_mean, d.gamma_rms, d.xi_mean, d.xi_rms, d.r_rms,
                          d.emit_nx, d.x_mean, d.y_mean, d.emit_ny, d.sigma_x, d.sigma_y, double(d.alive)};
    std::copy(v, v + 15, out15);
    return 0;
}
// J = gamma_rms / gamma_mean of beam b after one step, and dJ/dv of its weight profile (nseg values), by the
// reverse sweep of adjoint_oracle.inc; advances the run by that step
int or_sim_spread_grad(void* h, int b, double* J, double* grad, int nseg) {
    Sim& s = *static_cast<Sim*>(h);
    OR_TRY({
        if (b < 0 || b >= static_cast<int>(s.beams.size())) throw std::runtime_error("no such beam");
        if (nseg != static_cast<int>(s.beams[b].pwl.size())) throw std::runtime_error("profile size");
        adjoint::spread_grad(s, b, J, grad);
    })
}
// sets a test hook (0: beam sources, 1: fields) to n doubles (n = 0 clears it)
int or_sim_set_delta(void* h, int which, const double* d, long n) {
    Sim& s = *static_cast<Sim*>(h);
    std::vector<Real>& v = which == 0 ? s.bsrc_delta : s.fld_delta;
    const size_t want = which == 0 ? s.bsrc.size() : s.fld.size();
    if (n != 0 && static_cast<size_t>(n) != want) { last_error = "or_sim_set_delta: size"; return -1; }
    v.assign(d, d + n);
    return 0;
}
// after or_sim_spread_grad: the adjoints it formed (0: fields, 1: beam sources, 2: the beam's weights)
long or_sim_adjoint_get(void* h, int which, double* out, long cap) {
    Sim& s = *static_cast<Sim*>(h);
    const std::vector<double>& v = which == 0 ? s.adj_lf : which == 1 ? s.adj_lb : s.adj_lw;
    if (out && cap >= static_cast<long>(v.size())) std::copy(v.begin(), v.end(), out);
    return static_cast<long>(v.size());
}
// record chi_0 of every slice in the coming sweeps (or_sim_get "chi")
int or_sim_record_chi(void* h, int on) {
    Sim& s = *static_cast<Sim*>(h);
    s.record_chi = on != 0;
    s.chi_rec.assign(on ? static_cast<size_t>(s.box.nloc) * s.M : 0, 0.0);
    return 0;
}
int or_sim_write_fields(void* h, const char* path) { OR_TRY(write_fields(*static_cast<Sim*>(h), path)) }
int or_sim_write_axis(void* h, const char* path) { OR_TRY(write_axis(*static_cast<Sim*>(h), path)) }

}  // extern "C"
#include "laser_oracle.inc"
