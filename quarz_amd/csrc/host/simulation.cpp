#include "simulation.hpp"
#include "../qz_device.hpp"   // component order of fld / diag / bsrc

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <fstream>
#include <functional>
#include <iomanip>
#include <map>
#include <queue>
#include <random>
#include <sstream>
#include <stdexcept>
#include <thread>

namespace qz {

constexpr Real PI = 3.14159265358979323846;   // src/Types.hpp:22

// C-library pow taken at run time, never folded at build time, so that the host constants do not depend
// on the compiler's constant folding.
static double (*volatile c_pow)(double, double) = ::pow;

inline Real xi_from_user(Real xu) { return Real(0) - xu; }   // src/Types.hpp:30-31
inline Real xi_to_user(Real xi) { return Real(0) - xi; }

static void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
static void qz_check(int rc) {
    if (rc != QZ_OK) throw std::runtime_error(qz_last_error());
}

// ---------------------------------------------------------------------------- DevBuf
template <class T> void DevBuf<T>::alloc(size_t count) {
    release();
    n = count;
    if (count) hip_check(hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * count), "hipMalloc");
}
template <class T> void DevBuf<T>::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
}
template <class T> void DevBuf<T>::upload(const std::vector<T>& h, hipStream_t st) {
    if (h.size() > n) alloc(h.size());
    if (!h.empty()) hip_check(hipMemcpyAsync(p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, st), "upload");
}
template <class T> void DevBuf<T>::download(std::vector<T>& h, hipStream_t st) const {
    h.resize(n);
    if (n) hip_check(hipMemcpyAsync(h.data(), p, sizeof(T) * n, hipMemcpyDeviceToHost, st), "download");
    hip_check(hipStreamSynchronize(st), "download sync");   // the host reads h right after
}
template struct DevBuf<Real>;
template struct DevBuf<int>;
template struct DevBuf<char>;
template struct DevBuf<unsigned long long>;
template struct DevBuf<qz_sweep_args>;

// Stage n sweep argument blocks in the run's own device buffer and launch them on `st`.  The buffer is reused
// only by this run's later launches on the same stream; a larger batch first waits for the stream.
// With several workgroups per run the block map follows the argument blocks in the same buffer, and every run's
// barrier counter and error flag start at zero.
static void launch_sweeps(DevBuf<qz_sweep_args>& buf, const qz_sweep_args* host, int n, hipStream_t st) {
    bool multi = false;
    for (int i = 0; i < n; ++i) {
        if (qz_validate_sweep(host[i]) != QZ_OK) throw std::runtime_error(qz_last_error());
        multi = multi || host[i].groups > 1;
    }
    std::vector<int> map = multi ? qz_sweep_block_map(host, n) : std::vector<int>();
    const size_t map_blocks = (sizeof(int) * map.size() + sizeof(qz_sweep_args) - 1) / sizeof(qz_sweep_args);
    if (buf.n < static_cast<size_t>(n) + map_blocks) {
        hip_check(hipStreamSynchronize(st), "sync");
        buf.alloc(static_cast<size_t>(n) + map_blocks);
    }
    hip_check(hipMemcpyAsync(buf.p, host, sizeof(qz_sweep_args) * n, hipMemcpyHostToDevice, st), "sweep arguments");
    if (!multi) {
        qz_check(qz_sweep_launch(buf.p, host, n, st));
        return;
    }
    int* dmap = reinterpret_cast<int*>(buf.p + n);
    hip_check(hipMemcpy(dmap, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice), "block map");
    for (int i = 0; i < n; ++i) {
        if (!host[i].mws) throw std::runtime_error("a multi-workgroup launch needs every run's workspace");
        qz_check(qz_zero_async(host[i].mws, 64 * sizeof(double), st));
    }
    qz_check(qz_sweep_multi_launch(buf.p, dmap, host, n, st));
}

// a multi-workgroup sweep whose workgroups stopped waiting for each other (never in a correct launch) is an error
static void check_multi_sweep(const DevBuf<Real>& mws, int groups, hipStream_t st) {
    if (groups <= 1) return;
    unsigned int flag = 0;
    hip_check(hipMemcpyAsync(&flag, reinterpret_cast<const unsigned int*>(mws.p) + 2, sizeof(flag), hipMemcpyDeviceToHost,
                             st), "sweep flag");
    hip_check(hipStreamSynchronize(st), "sync");
    if (flag) throw std::runtime_error("multi-workgroup sweep: a grid barrier timed out (workgroups not co-resident)");
}

static int device_cus() {
    int dev = 0, ncu = 0;
    hip_check(hipGetDevice(&dev), "device");
    hip_check(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "device attribute");
    return ncu;
}

// Workgroups of one launch go to the accelerator dies (XCDs) in turn, block b to die b mod kXcds, so a launch of
// B workgroups puts at most ceil(B / kXcds) on any die, whichever die its first block goes to.  Two launches
// with W and B workgroups are resident together when ceil(W / kXcds) + ceil(B / kXcds) workgroups fit one die's
// compute units (one workgroup per compute unit).  MI355X: 8 dies of 32 compute units.
constexpr int kXcds = 8;
static int workers_beside(int ncu, int B) {
    return std::max(0, kXcds * (ncu / kXcds - (B + kXcds - 1) / kXcds));
}

static int mobile_particles(const Simulation& s) {
    int np = 0;
    for (const auto& sp : s.species)
        if (sp->mobile) np += sp->cap;
    return np;
}

// The sweeps of a batched step.  The one-workgroup runs are taken from a device queue, longest first (by
// Simulation::slice_ms when every run has one, else by slices x (cells + particles)), by workgroups that each
// hold a whole compute unit's LDS, each run swept with its own particles per thread (qz_sweep_mixed_launch).
// With multi-workgroup runs, when every one of them needs 2 particles per thread (the batch autotuner's choices
// always do) all runs share that one launch: its first
// `slots` blocks are the multi-workgroup runs' workgroups, packed so that each run's workgroups share a residue
// class mod kXcds and so one die and its L2, and the blocks the packing leaves empty, and the `workers` blocks
// after them, work the queue; at most one block per compute unit, so all are resident together.  Otherwise the
// multi-workgroup runs go to a launch of their own on a second stream, with the queue's workers kept to
// workers_beside(its workgroups) so that both launches are resident together whichever starts first.  Device
// layout of the arguments: the one-workgroup runs in queue order, the multi-workgroup runs, the block map, then
// one block whose first word is the queue counter.
struct MixedSweep {
    std::vector<qz_sweep_args> single, multi;
    std::vector<Simulation*> multi_runs;
    std::vector<int> map;
    std::vector<qz_sweep_args> stage;   // host image of the device arguments (alive until the step syncs)
    bool fused = false;                 // one launch for all runs
    int workers = 0, multi_blocks = 0, slots = 0;
    size_t map_offset() const { return single.size() + multi.size(); }
    size_t map_blocks() const { return (sizeof(int) * map.size() + sizeof(qz_sweep_args) - 1) / sizeof(qz_sweep_args); }
    size_t blocks() const { return map_offset() + map_blocks() + 1; }
};

static MixedSweep plan_mixed(Simulation* const* runs, int n, int max_workers) {
    MixedSweep P;
    bool est = true;
    for (int i = 0; i < n; ++i) est = est && runs[i]->slice_ms > 0;
    std::vector<std::pair<double, int>> cost;
    for (int i = 0; i < n; ++i) {
        Simulation& s = *runs[i];
        if (qz_validate_sweep(s.sweep_args) != QZ_OK) throw std::runtime_error(qz_last_error());
        if (s.groups > 1) {
            if (!s.mws.p) throw std::runtime_error("a multi-workgroup run needs its workspace");
            P.multi.push_back(s.sweep_args);
            P.multi_runs.push_back(&s);
            P.multi_blocks += s.groups;
        } else {
            cost.emplace_back(est ? s.slice_ms * s.K : static_cast<double>(s.K) * (s.grid.N + mobile_particles(s)), i);
        }
    }
    std::stable_sort(cost.begin(), cost.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
    for (const auto& c : cost) P.single.push_back(runs[c.second]->sweep_args);
    const int ncu = device_cus(), ns = static_cast<int>(P.single.size()), nm = static_cast<int>(P.multi.size());
    const int cap = max_workers > 0 ? max_workers : ncu;   // blocks that work the queue
    if (nm == 0) {
        P.fused = true;
        P.workers = std::min({ns, ncu, cap});
    } else {
        if (P.multi_blocks > ncu)
            throw std::runtime_error("batched sweep: the multi-workgroup runs need " + std::to_string(P.multi_blocks) +
                                     " workgroups, more than the " + std::to_string(ncu) + " compute units");
        P.fused = true;
        for (Simulation* r : P.multi_runs) P.fused = P.fused && qz_sweep_multi_pt(mobile_particles(*r), r->groups) == 2;
        if (P.fused) {
            std::vector<int> order(nm), load(kXcds, 0);
            for (int r = 0; r < nm; ++r) order[r] = r;
            std::stable_sort(order.begin(), order.end(),
                             [&](int x, int y) { return P.multi[x].groups > P.multi[y].groups; });
            std::vector<std::vector<std::pair<int, int>>> die(kXcds);
            for (int r : order) {
                const int x = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
                for (int g = 0; g < P.multi[r].groups; ++g) die[x].emplace_back(r, g);
                load[x] += P.multi[r].groups;
            }
            const int lmax = *std::max_element(load.begin(), load.end());
            std::vector<std::pair<int, int>> slot;   // (run, group); run -1: no run
            if (kXcds * lmax <= ncu) {
                slot.assign(kXcds * lmax, {-1, 0});
                for (int x = 0; x < kXcds; ++x)
                    for (size_t j = 0; j < die[x].size(); ++j) slot[x + kXcds * j] = die[x][j];
            } else {   // no room to pack: the runs' workgroups one after another
                for (int r = 0; r < nm; ++r)
                    for (int g = 0; g < P.multi[r].groups; ++g) slot.emplace_back(r, g);
            }
            P.slots = static_cast<int>(slot.size());
            int queue_blocks = 0;
            for (auto& sl : slot)
                if (sl.first < 0) {   // an empty slot works the queue (-1) or ends at once (-2)
                    const bool work = queue_blocks < std::min(ns, cap);
                    sl.first = work ? -1 : -2;
                    queue_blocks += work ? 1 : 0;
                }
            P.workers = std::max(0, std::min(std::min(ns, cap) - queue_blocks, ncu - P.slots));
            if (ns > 0 && queue_blocks + P.workers < 1)
                throw std::runtime_error("batched sweep: the multi-workgroup runs take every compute unit and leave "
                                         "none to the one-workgroup runs; give them fewer workgroups");
            P.map.resize(2 * slot.size());
            for (size_t b = 0; b < slot.size(); ++b) {
                P.map[2 * b] = slot[b].first;
                P.map[2 * b + 1] = slot[b].second;
            }
        } else {
            P.map = qz_sweep_block_map(P.multi.data(), nm);
            if (ns > 0 && workers_beside(ncu, P.multi_blocks) < 1)
                throw std::runtime_error("batched sweep: the multi-workgroup runs take every compute unit and leave "
                                         "none to the one-workgroup runs; give them fewer workgroups");
            P.workers = std::min({ns, workers_beside(ncu, P.multi_blocks), cap});
        }
    }
    P.stage.assign(P.blocks(), qz_sweep_args{});
    std::copy(P.single.begin(), P.single.end(), P.stage.begin());
    std::copy(P.multi.begin(), P.multi.end(), P.stage.begin() + static_cast<long>(P.single.size()));
    if (!P.map.empty()) std::memcpy(P.stage.data() + P.map_offset(), P.map.data(), sizeof(int) * P.map.size());
    return P;
}

// the arguments into `buf` (grown, after a wait for `st`, when too small), stream-ordered on st or at once
static void upload_mixed(const MixedSweep& P, DevBuf<qz_sweep_args>& buf, hipStream_t st, bool now) {
    if (buf.n < P.blocks()) {
        hip_check(hipStreamSynchronize(st), "sync");
        buf.alloc(P.blocks());
    }
    const size_t bytes = sizeof(qz_sweep_args) * (P.blocks() - 1);
    if (now) hip_check(hipMemcpy(buf.p, P.stage.data(), bytes, hipMemcpyHostToDevice), "sweep arguments");
    else hip_check(hipMemcpyAsync(buf.p, P.stage.data(), bytes, hipMemcpyHostToDevice, st), "sweep arguments");
}

// the launches (kernels only, so that a capture can hold them): one on `main`, or the multi launch forked to
// `aux` and joined back to `main` beside the queue's launch
static void enqueue_mixed(const MixedSweep& P, qz_sweep_args* dbuf, hipStream_t main, hipStream_t aux,
                          hipEvent_t fork, hipEvent_t join) {
    const int ns = static_cast<int>(P.single.size()), nm = static_cast<int>(P.multi.size());
    unsigned int* queue = reinterpret_cast<unsigned int*>(dbuf + (P.blocks() - 1));
    if (P.fused) {
        for (Simulation* s : P.multi_runs) qz_check(qz_zero_async(s->mws.p, 64 * sizeof(double), main));
        const int* dmap = reinterpret_cast<const int*>(dbuf + P.map_offset());
        qz_check(qz_sweep_mixed_launch(dbuf, P.single.data(), ns, ns > 0 ? queue : nullptr, dbuf + ns, dmap,
                                       P.multi.data(), nm, P.slots, P.workers, main));
        return;
    }
    if (nm > 0) {
        hip_check(hipEventRecord(fork, main), "fork");
        hip_check(hipStreamWaitEvent(aux, fork, 0), "fork");
        for (Simulation* s : P.multi_runs) qz_check(qz_zero_async(s->mws.p, 64 * sizeof(double), aux));
        const int* dmap = reinterpret_cast<const int*>(dbuf + P.map_offset());
        qz_check(qz_sweep_multi_launch(dbuf + ns, dmap, P.multi.data(), nm, aux, P.workers + P.multi_blocks));
        hip_check(hipEventRecord(join, aux), "join");
    }
    if (ns > 0)
        qz_check(qz_sweep_mixed_launch(dbuf, P.single.data(), ns, queue, nullptr, nullptr, nullptr, 0, 0, P.workers,
                                       main));
    if (nm > 0) hip_check(hipStreamWaitEvent(main, join, 0), "join");
}

// ---------------------------------------------------------------------------- RadialGrid (src/RadialGrid.cpp)
// The node lists of the three grid types.  Node positions, cell sizes and hat volumes are formed with the same
// floating-point operations as the reference's, so every array derived from them agrees bit for bit; the input
// checks and their messages are this library's.
namespace {
void grid_require(bool ok, const std::string& what) {
    if (!ok) throw std::runtime_error("radial grid: " + what);
}

// Spacing law of a graded grid: inside region i (up to its end radius) the cells aim at that region's spacing,
// and a finer region further out is approached so that the cells there arrive at its spacing, growing by at most
// the factor `ratio` per cell on the way (the target at x is at most h_i + (ratio - 1) * (start_i - x)).
struct Grading {
    std::vector<std::pair<double, double>> regions;   // (end radius, spacing), end radii increasing
    Real grow;                                        // ratio - 1
    Real target(Real x) const {
        Real h = regions.back().second;
        for (const auto& reg : regions)
            if (x < reg.first) {
                h = reg.second;
                break;
            }
        Real start = 0;   // region i starts where region i - 1 ends
        for (const auto& reg : regions) {
            if (start > x) h = std::min(h, grow > 0 ? Real(reg.second + grow * (start - x)) : Real(reg.second));
            start = reg.first;
        }
        return h;
    }
};
}  // namespace

std::vector<Real> RadialGrid::make_uniform(Real R, Real dr) {   // :14-20
    grid_require(R > 0, "grid.rmax must be positive");
    grid_require(dr > 0, "grid.dr must be positive");
    const int cells = std::max(1, static_cast<int>(std::lround(R / dr)));
    std::vector<Real> nodes;
    nodes.reserve(static_cast<size_t>(cells) + 1);
    for (int j = 0; j <= cells; ++j) nodes.push_back(R * Real(j) / Real(cells));
    return nodes;
}

// March outwards cell by cell: each new cell takes the target spacing at its middle (estimated with the previous
// cell's size), limited to a factor `ratio` from the previous cell; the first cell takes the target at the axis.
// The last node is dropped when it overshoots R by more than the one before falls short, and the nodes are scaled
// to end exactly at R.
std::vector<Real> RadialGrid::make_regions(Real R, const std::vector<std::pair<double, double>>& regions,
                                           Real max_ratio) {   // :27-71
    grid_require(!regions.empty(), "grid.regions lists no region");
    grid_require(max_ratio >= 1.0, "grid.max_ratio below 1 would shrink every cell");
    for (size_t i = 0; i < regions.size(); ++i) {
        grid_require(regions[i].second > 0, "every region spacing of grid.regions must be positive");
        grid_require(i == 0 || regions[i].first > regions[i - 1].first,
                     "the end radii of grid.regions must increase from region to region");
    }
    const Grading law{regions, max_ratio - 1.0};
    std::vector<Real> nodes{0.0};
    Real x = 0, prev = law.target(0);
    while (x < R * (1 - 1e-12)) {
        Real h = std::clamp(law.target(x + 0.5 * prev), prev / max_ratio, prev * max_ratio);
        if (nodes.size() == 1) h = law.target(0);
        x += h;
        nodes.push_back(x);
        prev = h;
        grid_require(nodes.size() <= 10000000, "grid.regions asks for more than 10^7 cells");
    }
    const size_t n = nodes.size();
    if (n >= 3 && nodes[n - 1] - R > R - nodes[n - 2]) nodes.pop_back();
    const Real scale = R / nodes.back();
    for (Real& v : nodes) v *= scale;
    nodes.back() = R;
    return nodes;
}

// one node per line; '#' starts a comment; blank lines are skipped
std::vector<Real> RadialGrid::read_file(const std::string& filename) {   // :73-85
    std::ifstream in(filename);
    grid_require(static_cast<bool>(in), "grid.file '" + filename + "' cannot be opened");
    std::vector<Real> nodes;
    int lineno = 0;
    for (std::string line; std::getline(in, line);) {
        ++lineno;
        line = line.substr(0, line.find('#'));
        const char* p = line.c_str();
        char* end = nullptr;
        const double v = std::strtod(p, &end);
        if (end == p) {
            grid_require(line.find_first_not_of(" \t\r") == std::string::npos,
                         "grid.file '" + filename + "' line " + std::to_string(lineno) + " is not a radius");
            continue;
        }
        nodes.push_back(v);
    }
    return nodes;
}

void RadialGrid::finalize(const std::vector<Real>& nodes, hipStream_t st) {   // :106-148
    grid_require(nodes.size() >= 3, "a grid needs at least two cells");
    grid_require(nodes[0] == 0, "the first node must sit on the axis (r = 0)");
    for (size_t j = 1; j < nodes.size(); ++j)
        grid_require(nodes[j] > nodes[j - 1], "the nodes must increase strictly (node " + std::to_string(j) + ")");
    r = nodes;
    N = static_cast<int>(r.size()) - 1;
    R = r[N];
    h.assign(N, 0.0);
    rmid.assign(N, 0.0);
    inv_h.assign(N, 0.0);
    // hat volume of node j, 2 pi Int W_j r dr: cell c contributes r_c h/2 + h^2/6 to its inner node and
    // r_c h/2 + h^2/3 to its outer node (the inner part first, as the reference adds them)
    std::vector<Real> half(N + 1, 0.0);
    for (int c = 0; c < N; ++c) {
        h[c] = r[c + 1] - r[c];
        rmid[c] = 0.5 * (r[c + 1] + r[c]);
        inv_h[c] = 1.0 / h[c];
        half[c + 1] += r[c] * h[c] / 2 + h[c] * h[c] / 3;
    }
    V.assign(N + 1, 0.0);
    inv_V.assign(N + 1, 0.0);
    for (int j = 0; j <= N; ++j) {
        Real v = half[j];
        if (j < N) v += r[j] * h[j] / 2 + h[j] * h[j] / 6;
        V[j] = 2 * PI * v;
        inv_V[j] = 1.0 / V[j];
    }
    d_r.upload(r, st);
    d_inv_h.upload(inv_h, st);
    d_inv_V.upload(inv_V, st);
}

// ---------------------------------------------------------------------------- FieldSolver (src/FieldSolver.cpp)
namespace {
struct Rows {
    std::vector<Real> a, b, c;
    explicit Rows(int M) : a(M, 0.0), b(M, 0.0), c(M, 0.0) {}
};

// L0: the lumped finite-element rows of the axisymmetric Laplacian, a Dirichlet row at the wall (:43-59)
Rows laplacian_m0(const RadialGrid& g) {
    const int N = g.N;
    Rows L(N + 1);
    for (int j = 0; j < N; ++j) {
        const Real scale = 2 * PI / g.V[j];
        L.a[j] = j > 0 ? g.rmid[j - 1] / g.h[j - 1] * scale : 0.0;
        L.c[j] = g.rmid[j] / g.h[j] * scale;
        L.b[j] = -(L.a[j] + L.c[j]);
    }
    L.b[N] = 1.0;
    return L;
}

// L_n, n >= 1: finite-volume rows of r^(n-1) d/dr (r^(1-2n) d/dr (r^n f)) / r^... written for f = r^n g with the
// flux of cell j weighted by 1 / (h_j rmid_j^(2n-1)) (:20-33, 63-85); Dirichlet rows at the axis and, unless the
// wall row is the flux (B_theta) condition, at the wall.  flux_coef receives the wall row's source factor.
Rows laplacian_mn(const RadialGrid& g, int n, bool flux_wall, double* flux_coef) {
    const int N = g.N;
    Rows L(N + 1);
    auto kp = [&](int j) { return 1.0 / (g.h[j] * c_pow(g.rmid[j], 2 * n - 1)); };
    L.b[0] = 1.0;
    for (int j = 1; j < N; ++j) {
        const Real dV = 0.5 * (g.h[j] + g.h[j - 1]);
        const Real pre = c_pow(g.r[j], n - 1) / dV;
        L.c[j] = pre * c_pow(g.r[j + 1], n) * kp(j);
        L.a[j] = pre * c_pow(g.r[j - 1], n) * kp(j - 1);
        L.b[j] = -pre * c_pow(g.r[j], n) * (kp(j) + kp(j - 1));
    }
    if (flux_wall) {
        const Real dV = 0.5 * g.h[N - 1];
        const Real k = kp(N - 1) / dV;
        L.a[N] = g.r[N - 1] * k;
        L.b[N] = -g.r[N] * k;
        *flux_coef = 1.0 / dV;
    } else {
        L.b[N] = 1.0;
    }
    return L;
}

// Thomas factors c' and m of a chi-free operator, formed as TridiagSolver's loop forms them (src/Tridiag.cpp:114-121)
void thomas_factors(const Rows& L, std::vector<Real>& m, std::vector<Real>& cp) {
    const size_t M = L.b.size();
    m.assign(M, 0.0);
    cp.assign(M, 0.0);
    cp[0] = L.c[0] / L.b[0];
    for (size_t i = 1; i < M; ++i) {
        m[i] = Real(1) / (L.b[i] - L.a[i] * cp[i - 1]);
        cp[i] = L.c[i] * m[i];
    }
}
}  // namespace

void FieldSolver::init(const RadialGrid& g, hipStream_t st) {   // :43-109, 157-176
    const int N = g.N, M = N + 1;
    const Rows ops[4] = {laplacian_m0(g), laplacian_mn(g, 1, false, nullptr), laplacian_mn(g, 1, true, &flux_coef),
                         laplacian_mn(g, 2, false, nullptr)};   // order L0, L1D, L1F, L2D
    for (int op = 0; op < 4; ++op) {
        a[op] = ops[op].a;
        b[op] = ops[op].b;
        c[op] = ops[op].c;
    }
    // three-point derivative on the non-uniform grid (interior: centred; wall: one-sided), and 1/r with the
    // axis value 1/r_1
    gm.assign(M, 0); g0.assign(M, 0); gp.assign(M, 0); ir.assign(M, 0);
    for (int j = 1; j < N; ++j) {
        const Real hm = g.h[j - 1], hp = g.h[j];
        const Real den = hm * hp * (hm + hp);
        gp[j] = hm * hm / den;
        gm[j] = -hp * hp / den;
        g0[j] = (hp * hp - hm * hm) / den;
    }
    const Real h1 = g.h[N - 1], h2 = N >= 2 ? g.h[N - 2] : g.h[N - 1];
    const Real s12 = h1 + h2;
    g0[N] = (2 * h1 + h2) / (h1 * s12);
    gm[N] = -s12 / (h1 * h2);
    gp[N] = h1 / (h2 * s12);
    for (int j = 1; j <= N; ++j) ir[j] = 1.0 / g.r[j];
    ir[0] = 1.0 / g.r[1];
    // the chi-free solves (psi, E_z, B_z) keep b as it is: factors formed once per run
    thomas_factors(ops[L0], l0_m, l0_cp);
    thomas_factors(ops[L1D], l1_m, l1_cp);
    for (int op = 0; op < 4; ++op) {
        d_a[op].upload(a[op], st); d_b[op].upload(b[op], st); d_c[op].upload(c[op], st);
    }
    d_gm.upload(gm, st); d_g0.upload(g0, st); d_gp.upload(gp, st); d_ir.upload(ir, st);
    d_l0_m.upload(l0_m, st); d_l0_cp.upload(l0_cp, st); d_l1_m.upload(l1_m, st); d_l1_cp.upload(l1_cp, st);
}

// ---------------------------------------------------------------------------- PlasmaSpecies (src/PlasmaSpecies.cpp)
// Adams-Bashforth weights of orders 1 to 5 for uniform steps, newest force first (the method's table; the
// reference's make_ab, :14-28, holds the same values)
static void make_ab(int order, Real* c) {
    static const Real table[5][5] = {
        {1, 0, 0, 0, 0},
        {1.5, -0.5, 0, 0, 0},
        {23. / 12, -16. / 12, 5. / 12, 0, 0},
        {55. / 24, -59. / 24, 37. / 24, -9. / 24, 0},
        {1901. / 720, -2774. / 720, 2616. / 720, -1274. / 720, 251. / 720},
    };
    if (order < 1 || order > 5) throw std::runtime_error("pusher.ab_order: Adams-Bashforth orders 1 to 5 exist, not " +
                                                         std::to_string(order));
    for (int i = 0; i < 5; ++i) c[i] = table[order - 1][i];
}

void PlasmaSpecies::init(const Config& cfg, const std::string& nm, const RadialGrid& g, bool mode1) {   // :92-185
    name = nm;
    m1 = mode1;
    auto key = [&](const char* k) { return nm + "." + k; };
    q = cfg.get_double(key("charge"), -1.0);
    m = cfg.get_double(key("mass"), 1.0);
    mobile = cfg.get_bool(key("mobile"), true);
    ppc = cfg.get_int(key("ppc"), 4);
    ntheta = cfg.get_int(key("ntheta"), m1 ? 8 : 1);
    uth = cfg.get_double(key("uth"), 0.0);
    density = cfg.get_double(key("density"), 1.0);
    rmax = cfg.get_double("plasma.rmax", 1e300);
    delta_min = cfg.get_double("pusher.delta_min", 1e-3);
    max_qsa = cfg.get_double("pusher.max_qsa_factor", 35.0);
    ab_order = cfg.get_int("pusher.ab_order", 3);
    const std::string profile = cfg.get_string("plasma.radial", "uniform");
    if (profile != "uniform")
        throw std::runtime_error("plasma.radial = " + profile + ": this step loads the uniform profile only");
    if (ppc < 1) throw std::runtime_error(key("ppc") + " must be at least 1 ring per cell");
    if (ntheta < 1) throw std::runtime_error(key("ntheta") + " must be at least 1 particle per ring");
    if (m1 && ntheta < 4)
        throw std::runtime_error(key("ntheta") + " must be at least 4 with azimuthal mode 1 (8 is the default)");
    make_ab(ab_order, ab_c);
    cap = std::max(g.N * ppc * ntheta, 1);
    Np = cap;
    for (DevBuf<Real>* a : {&x, &y, &px, &py, &dl, &w}) a->alloc(cap);
    hist.alloc(static_cast<size_t>(ab_order) * 5 * cap);
}

// The rings at the head of the box (src/PlasmaSpecies.cpp:202-295): every cell holds ppc rings of equal radial
// width, each at the volume centroid r = (2/3)(rb^3 - ra^3) / (rb^2 - ra^2) of its annulus [ra, rb], with the
// annulus volume pi (rb^2 - ra^2) per unit length times the density as its weight (zero beyond plasma.rmax);
// every other ring is turned by half an azimuth step, and a ring is ntheta particles at equal angles (one on
// the x axis when ntheta = 1).  A thermal plasma draws ux, uy, uz per particle from std::normal_distribution in
// that order.  Particles of zero weight are not loaded.
void PlasmaSpecies::load(const RadialGrid& g, Real z, unsigned long long seed, hipStream_t st) {
    (void)z;   // the longitudinal profile is flat on this path
    if (device_rings()) return;   // the sweep loads the rings and clears the history itself
    if (have_template) {   // the rings of the first load (device copies: capturable, no host work)
        const size_t bytes = sizeof(Real) * static_cast<size_t>(Np);
        const DevBuf<Real>* from[6] = {&t_x, &t_y, &t_px, &t_py, &t_dl, &t_w};
        DevBuf<Real>* to[6] = {&x, &y, &px, &py, &dl, &w};
        for (int a = 0; a < 6 && bytes; ++a)
            hip_check(hipMemcpyAsync(to[a]->p, from[a]->p, bytes, hipMemcpyDeviceToDevice, st), "plasma copy");
        qz_check(qz_zero_async(hist.p, sizeof(Real) * hist.n, st));
        return;
    }
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> gauss(0.0, 1.0);
    for (std::vector<Real>* a : {&hx, &hy, &hpx, &hpy, &hdl, &hw}) a->assign(cap, 0.0);
    int n = 0;
    for (int ring = 0; ring < g.N * ppc; ++ring) {
        const int j = ring / ppc, sub = ring % ppc;
        const double hc = g.h[j];
        const double ra = g.r[j] + sub * hc / ppc, rb = g.r[j] + (sub + 1) * hc / ppc;
        const double rc = (2.0 / 3.0) * (rb * rb * rb - ra * ra * ra) / (rb * rb - ra * ra);
        const double volume = (1.0 * density) * PI * (rb * rb - ra * ra);
        const double wp = (rc > rmax ? 0.0 : 1.0) * volume / ntheta;
        const double turn = (ring % 2) * 0.5;
        for (int l = 0; l < ntheta; ++l) {
            double u[3] = {0, 0, 0};
            if (uth > 0)
                for (double& v : u) v = uth * gauss(rng);
            if (wp == 0) continue;
            const double th = 2 * PI * (l + turn) / ntheta;
            const double gam = std::sqrt(1 + u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            hx[n] = ntheta == 1 ? rc : rc * std::cos(th);
            hy[n] = ntheta == 1 ? 0.0 : rc * std::sin(th);
            hpx[n] = u[0];
            hpy[n] = u[1];
            hdl[n] = gam - u[2];
            hw[n] = wp;
            ++n;
        }
    }
    Np = n;
    const std::vector<Real>* host[6] = {&hx, &hy, &hpx, &hpy, &hdl, &hw};
    DevBuf<Real>* live[6] = {&x, &y, &px, &py, &dl, &w};
    DevBuf<Real>* keep[6] = {&t_x, &t_y, &t_px, &t_py, &t_dl, &t_w};
    for (int a = 0; a < 6; ++a) {
        live[a]->upload(*host[a], st);
        keep[a]->alloc(cap);
        keep[a]->upload(*host[a], st);
    }
    have_template = true;
    qz_check(qz_zero_async(hist.p, sizeof(Real) * hist.n, st));
    // whole rings are loaded or skipped, and the skipped ones lie beyond the loaded ones (radius grows with ring)
    nring = Np / ntheta;
    std::vector<Real> cs(4 * static_cast<size_t>(ntheta));
    for (int t = 0; t < 2; ++t)
        for (int l = 0; l < ntheta; ++l) {
            const double th = 2 * PI * (l + t * 0.5) / ntheta;
            cs[t * ntheta + l] = std::cos(th);
            cs[2 * ntheta + t * ntheta + l] = std::sin(th);
        }
    d_cs.upload(cs, st);
}

// host deposit of the loaded rings (background / neutralization), serial index order
void PlasmaSpecies::deposit_rho(const RadialGrid& g, std::vector<Real>& rhot, std::vector<Real>& rho) const {
    const int M = g.N + 1;
    for (int i = 0; i < Np; ++i) {
        const Real wi = hw[i];
        if (wi == Real(0)) continue;
        const Real qq = q * Real(1);
        if (qq == Real(0)) continue;
        const Real X = hx[i], Y = hy[i];
        const Real r = std::sqrt(X * X + Y * Y);
        Real c, s;
        if (r > Real(0)) { c = X / r; s = Y / r; } else { c = Real(1); s = Real(0); }
        int j;
        if (r <= Real(0)) j = 0;
        else if (r >= g.R) j = g.N - 1;
        else {
            int lo = 0, hi = g.N;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (g.r[mid] <= r) lo = mid; else hi = mid;
            }
            j = lo;
        }
        Real t = (r - g.r[j]) * g.inv_h[j];
        t = t < Real(0) ? Real(0) : (t > Real(1) ? Real(1) : t);
        const Real W[2] = {Real(1) - t, t};
        for (int cc = 0; cc < 2; ++cc) {
            const int jj = j + cc;
            const Real a = wi * qq * W[cc] * g.inv_V[jj];
            rhot[jj] += a;
            rho[jj] += a;
            if (m1) {
                rhot[M + jj] += a * c;
                rhot[2 * M + jj] += -a * s;
                rho[M + jj] += a * c;
                rho[2 * M + jj] += -a * s;
            }
        }
    }
}

// ---------------------------------------------------------------------------- Beam (src/Beam.cpp)
namespace {
// What a deck says about one particle beam of this path (src/Beam.cpp:53-207, 483-570): the Gaussian or
// flat-top bunch, its transverse offset and tilt, its momenta, its symmetry copies and the optional
// piecewise-linear current profile of a flat-top witness.
struct BeamSpec {
    bool gaussian = true;
    double n0 = 0, sr = 0, xi0 = 0, sxi = 0, len = 0, gamma0 = 0, espread = 0, emit = 0, cut = 3;
    double cx0 = 0, cy0 = 0, cxs = 0, cys = 0, xp0 = 0, yp0 = 0, xcut = -1e300;
    int ntot = 0, nrot = 1, nsym = 1;
    unsigned long long seed = 12345;
    std::vector<double> pwl;
};

BeamSpec read_beam_spec(const Config& cfg, const std::string& nm, bool m1) {
    auto key = [&](const char* k) { return nm + "." + k; };
    BeamSpec b;
    const std::string profile = cfg.get_string(key("profile"), "gaussian");
    if (profile != "gaussian" && profile != "flattop")
        throw std::runtime_error(key("profile") + " = " + profile + ": this step samples gaussian and flattop beams");
    b.gaussian = profile == "gaussian";
    b.n0 = cfg.get_double(key("n0"));
    b.sr = cfg.get_double(key("sigma_r"));
    b.xi0 = xi_from_user(cfg.get_double(key("xi0")));
    b.gamma0 = cfg.get_double(key("gamma"));
    b.espread = cfg.get_double(key("espread"), 0.0);
    b.emit = cfg.get_double(key("emit_n"), 0.0);
    b.cut = cfg.get_double(key("cut"), 3.0);
    b.cx0 = cfg.get_double(key("x0"), 0.0);
    b.cy0 = cfg.get_double(key("y0"), 0.0);
    b.cxs = -cfg.get_double(key("x_slope"), 0.0);   // d x_c / d xi = - d x_c / d xi_user
    b.cys = -cfg.get_double(key("y_slope"), 0.0);
    b.xp0 = cfg.get_double(key("xp0"), 0.0);
    b.yp0 = cfg.get_double(key("yp0"), 0.0);
    b.ntot = cfg.get_int(key("nparticles"), 100000);
    b.nrot = cfg.get_int(key("nsym"), 1);
    b.nsym = b.nrot * (cfg.get_bool(key("mirror_y"), false) ? 2 : 1);
    b.seed = cfg.get_int(key("seed"), 12345);
    if (cfg.has(key("xi_cut_head"))) b.xcut = xi_from_user(cfg.get_double(key("xi_cut_head")));
    if (b.gaussian) b.sxi = cfg.get_double(key("sigma_xi"));
    else b.len = cfg.get_double(key("length"));
    for (const auto& tok : cfg.get_list(key("pwl_weights"))) b.pwl.push_back(cfg.number(tok, key("pwl_weights")));
    const bool offset = b.cx0 != 0 || b.cy0 != 0 || b.cxs != 0 || b.cys != 0 || b.xp0 != 0 || b.yp0 != 0;
    if (offset && !m1) throw std::runtime_error(nm + ": an off-axis or tilted beam needs modes = 1");
    if (b.nrot < 1) throw std::runtime_error(key("nsym") + " must be at least 1");
    if (b.ntot % b.nsym != 0)
        throw std::runtime_error(key("nparticles") + " = " + std::to_string(b.ntot) + " is not a multiple of the " +
                                 std::to_string(b.nsym) + " symmetry copies (nsym, doubled by mirror_y)");
    if (!b.pwl.empty() && (b.gaussian || b.nsym != 1))
        throw std::runtime_error(key("pwl_weights") + ": a current profile needs a flattop beam without symmetry copies");
    return b;
}

// Weight of one macro-particle: the bunch charge in units of n0 (c/omega_p)^3 inside the radial cut and the
// longitudinal window (after the head cut), shared by the macro-particles.
double macro_weight(const BeamSpec& b) {
    const double inside_r = 1.0 - std::exp(-0.5 * b.cut * b.cut);   // fraction inside r < cut sigma_r
    double total;
    if (b.gaussian) {
        auto Phi = [](double v) { return 0.5 * (1.0 + std::erf(v / std::sqrt(2.0))); };
        const double lo = std::max(-b.cut, (b.xcut - b.xi0) / b.sxi);
        if (lo >= b.cut) throw std::runtime_error("xi_cut_head lies behind the whole bunch");
        total = b.n0 * c_pow(2 * PI, 1.5) * b.sr * b.sr * b.sxi * inside_r * (Phi(b.cut) - Phi(lo));
    } else {
        const double lo = std::max(b.xi0 - 0.5 * b.len, b.xcut), hi = b.xi0 + 0.5 * b.len;
        if (lo >= hi) throw std::runtime_error("xi_cut_head lies behind the whole bunch");
        total = b.n0 * 2 * PI * b.sr * b.sr * inside_r * (hi - lo);
    }
    return total / b.ntot;
}
}  // namespace

void Beam::init(const Config& cfg, const std::string& nm, bool mode1, hipStream_t st) {
    name = nm;
    m1 = mode1;
    q = cfg.get_double(nm + ".charge", -1.0);
    m = cfg.get_double(nm + ".mass", 1.0);
    rigid = cfg.get_bool(nm + ".rigid", false);
    static const std::map<std::string, int> pushers = {{"vay", 0}, {"boris", 1}, {"hc", 2}, {"imp", 3}};
    const std::string pm = cfg.get_string(nm + ".pusher", cfg.get_string("pusher.beam", "vay"));
    if (!pushers.count(pm)) throw std::runtime_error(nm + ": pusher '" + pm + "' is not one of vay, boris, hc, imp");
    pusher = pushers.at(pm);
    const BeamSpec b = read_beam_spec(cfg, nm, m1);
    const double wmac = macro_weight(b);
    pwl = b.pwl;
    xi_init.clear();
    g_init.clear();
    if (!pwl.empty()) {   // the profile runs over the flat top, tail first, in user xi
        pwl_u0 = xi_to_user(b.xi0 + 0.5 * b.len);
        pwl_u1 = xi_to_user(b.xi0 - 0.5 * b.len);
        pwl_wmac = wmac;
    }
    // Sampling, one draw sequence of std::mt19937_64(seed) per group of nsym copies: (x, y) from the radial
    // Gaussian inside cut sigma_r; xi from the longitudinal shape, redrawn ahead of the head cut; px, py from
    // the emittance; gamma from the energy spread.  The copies are the sample turned by 2 pi l / nrot (and
    // mirrored in y), offset and tilted by the centroid, with the momenta turned alike.
    std::mt19937_64 rng(b.seed);
    std::normal_distribution<double> N01(0.0, 1.0);
    std::uniform_real_distribution<double> U01(0.0, 1.0);
    const double spx = b.sr > 0 ? b.emit / b.sr : 0.0;
    double aw = 0, awg = 0, awxi = 0;
    std::vector<double> rows;
    rows.reserve(static_cast<size_t>(b.ntot) * 7);
    for (int group = 0; group < b.ntot / b.nsym; ++group) {
        double xx, yy;
        do {
            xx = b.sr * N01(rng);
            yy = b.sr * N01(rng);
        } while (xx * xx + yy * yy > b.cut * b.cut * b.sr * b.sr);
        double xv;
        do {
            if (b.gaussian) {
                do xv = N01(rng); while (std::abs(xv) > b.cut);
                xv = b.xi0 + b.sxi * xv;
            } else {
                xv = b.xi0 + b.len * (U01(rng) - 0.5);
            }
        } while (xv < b.xcut);
        double ux = spx * N01(rng), uy = spx * N01(rng);
        const double gam = b.gamma0 * (1.0 + b.espread * N01(rng));
        ux += gam * b.xp0;
        uy += gam * b.yp0;
        const double pz2 = gam * gam - 1.0 - ux * ux - uy * uy;
        if (pz2 <= 0) throw std::runtime_error(nm + ": the transverse momentum spread exceeds what gamma allows");
        const double tpx = ux - gam * b.xp0, tpy = uy - gam * b.yp0;
        const double shift = xv - b.xi0;
        for (int l = 0; l < b.nsym; ++l) {
            const double mirror = l >= b.nrot ? -1.0 : 1.0;
            const double ph = 2 * PI * (l % b.nrot) / b.nrot, cp = std::cos(ph), sp = std::sin(ph);
            const double wgt = pwl.empty() ? wmac : wmac * pwl_factor(pwl, pwl_u0, pwl_u1, xi_to_user(xv));
            const double row[7] = {cp * xx - sp * yy + b.cx0 + b.cxs * shift,
                                   mirror * (sp * xx + cp * yy) + b.cy0 + b.cys * shift,
                                   cp * tpx - sp * tpy + gam * b.xp0,
                                   mirror * (sp * tpx + cp * tpy) + gam * b.yp0,
                                   std::sqrt(pz2), xv, wgt};
            if (wgt != 0 || !pwl.empty()) {   // the reference values of the moments: mean gamma and mean xi
                const double gg = std::sqrt(1 + row[2] * row[2] + row[3] * row[3] + row[4] * row[4]);
                if (!pwl.empty()) {
                    xi_init.push_back(xv);
                    g_init.push_back(gg);
                }
                aw += wgt;
                awg += wgt * gg;
                awxi += wgt * xv;
            }
            rows.insert(rows.end(), row, row + 7);
        }
    }
    set_particles(rows, st);
    gref = aw > 0 ? awg / aw : 0;
    xiref = aw > 0 ? awxi / aw : 0;
}

// The profile's values sit at the centres of nseg equal segments of [u0, u1] (user xi, tail first); the
// profile is linear between neighbouring centres and constant beyond the first and the last centre.
double pwl_factor(const std::vector<double>& v, double u0, double u1, double u) {
    const int S = static_cast<int>(v.size());
    const double t = (u - u0) / (u1 - u0) * S - 0.5;
    if (!(t > 0)) return v[0];
    if (t >= S - 1) return v[S - 1];
    const int sg = static_cast<int>(t);
    const double f = t - sg;
    return v[sg] + (v[sg + 1] - v[sg]) * f;
}

void Beam::set_profile(const std::vector<double>& v, hipStream_t st) {
    if (pwl.empty()) throw std::runtime_error(name + ": no piecewise-linear profile (pwl_weights)");
    if (v.size() != pwl.size()) throw std::runtime_error(name + ": profile needs " + std::to_string(pwl.size()) + " values");
    for (double x : v)
        if (!(x > 0) || !std::isfinite(x)) throw std::runtime_error(name + ": profile values must be positive and finite");
    pwl = v;
    std::vector<double> w(xi_init.size());
    double aw = 0, awg = 0, awxi = 0;   // the moments' reference values, as init makes them
    for (size_t i = 0; i < w.size(); ++i) {
        w[i] = pwl_wmac * pwl_factor(pwl, pwl_u0, pwl_u1, xi_to_user(xi_init[i]));
        if (w[i] != 0) {
            aw += w[i];
            awg += w[i] * g_init[i];
            awxi += w[i] * xi_init[i];
        }
    }
    gref = 0; xiref = 0;
    if (aw > 0) { gref = awg / aw; xiref = awxi / aw; }
    a[6].upload(w, st);
    hip_check(hipStreamSynchronize(st), "sync");
}

void Beam::set_particles(const std::vector<double>& p7, hipStream_t st) {   // :1323-1335
    Np = static_cast<int>(p7.size() / 7);
    std::vector<double> col(static_cast<size_t>(Np));
    for (int c = 0; c < 7; ++c) {
        for (int i = 0; i < Np; ++i) col[i] = p7[7 * static_cast<size_t>(i) + c];
        a[c].alloc(static_cast<size_t>(std::max(Np, 1)));
        if (Np) hip_check(hipMemcpyAsync(a[c].p, col.data(), 8 * static_cast<size_t>(Np), hipMemcpyHostToDevice, st), "beam upload");
        hip_check(hipStreamSynchronize(st), "sync");
    }
}

std::array<double, 17> Beam::sums(hipStream_t st) const {
    std::array<double, 17> S{};
    const qz_beam_soa s = soa();
    qz_check(qz_beam_sums(Np, &s, gref, xiref, S.data(), st));
    return S;
}

std::array<double, 17> Beam::sums_device(hipStream_t st) const {
    const size_t need = 17 + qz_beam_sums_workspace_bytes(Np) / sizeof(Real);
    if (sums_work.n < need) sums_work.alloc(need);
    const qz_beam_soa s = soa();
    qz_check(qz_beam_sums_device(Np, &s, gref, xiref, sums_work.p, sums_work.p + 17, st));
    std::array<double, 17> S{};
    hip_check(hipMemcpyAsync(S.data(), sums_work.p, sizeof(S), hipMemcpyDeviceToHost, st), "sums");
    hip_check(hipStreamSynchronize(st), "sync");
    return S;
}

BeamDiag Beam::from_sums(const std::array<double, 17>& S, Real t) const {   // :1157-1192
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
        if (!m1) {   // m = 0: ring averages over the azimuth
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

// keys of this path: full names and the per-species / per-beam suffixes (the known-key rule of src/main.cpp:26-37)
static bool known_key(const std::string& key) {
    static const char* const full[] = {
        "xi.convention", "modes", "solver.picard", "solver.tridiag", "plasma.smooth_length", "pusher.max_cells_per_step",
        "grid.type", "grid.rmax", "grid.dr", "grid.regions", "grid.max_ratio", "grid.file", "xi.max", "xi.min",
        "xi.step", "time.dt", "time.steps", "time.start", "time.first_step", "plasma.seed", "beams.xi_shape",
        "plasma.species", "plasma.neutralize", "plasma.rmax", "plasma.radial", "beams", "pusher.beam",
        "pusher.ab_order", "pusher.delta_min", "pusher.max_qsa_factor", "output.dir", "output.every", "sweep.groups",
        "output.beam_every", "input.allow_unknown", "units.n0", "units.n0_cm3"};
    static const char* const suffix[] = {
        ".charge", ".mass", ".mobile", ".ppc", ".ntheta", ".uth", ".density", ".rigid", ".pusher", ".profile", ".n0",
        ".sigma_r", ".xi0", ".gamma", ".espread", ".emit_n", ".cut", ".x0", ".y0", ".x_slope", ".y_slope", ".xp0",
        ".yp0", ".nparticles", ".nsym", ".mirror_y", ".seed", ".xi_cut_head", ".sigma_xi", ".length"};
    for (const char* f : full)
        if (key == f) return true;
    for (const char* s : suffix) {
        const std::string sx(s);
        if (key.size() > sx.size() && key.compare(key.size() - sx.size(), sx.size(), sx) == 0) return true;
    }
    return false;
}

// ---------------------------------------------------------------------------- Simulation (src/Simulation.cpp)
Simulation::Simulation(const std::string& deck, const std::string& overrides) {   // :34-318 (the keys of this path)
    Config cfg;
    cfg.parse(deck);
    cfg.overrides(overrides);
    // a laser (src/Simulation.cpp:272) pushes the plasma and the beams ponderomotively; this sweep and these pushes
    // have no such term, so such a deck is refused here rather than run without it (the envelope alone: qz_laser_*)
    if (cfg.has("laser.a0") || cfg.has("laser.file") || cfg.has("laser.comb_a"))
        throw std::runtime_error("laser: this step has no ponderomotive coupling, so decks with a laser are not run "
                                 "(the laser envelope alone is the qz_laser_* component)");
    hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    for (auto& e : ev) hip_check(hipEventCreate(&e), "hipEventCreate");
    if (!cfg.has("xi.convention"))
        throw std::runtime_error("xi.convention is missing (QUARZ v2 decks state xi.convention = z-t)");
    if (cfg.get_string("xi.convention") != "z-t") throw std::runtime_error("xi.convention: only z-t is supported");
    const int modes = cfg.get_int("modes", 0);
    if (modes != 0 && modes != 1) throw std::runtime_error("modes must be 0 (axisymmetric) or 1 (m = 0 and 1)");
    m1 = (modes == 1);
    picard = cfg.get_int("solver.picard", m1 ? 3 : 1);
    (void)cfg.get_string("solver.tridiag", "auto");   // every method solves the same system; Thomas order here
    if (cfg.get_double("plasma.smooth_length", 0.0) != 0.0) throw std::runtime_error("plasma.smooth_length is not on this path");
    if (cfg.get_double("pusher.max_cells_per_step", 0.0) != 0.0) throw std::runtime_error("sub-slicing is not on this path");
    {
        const std::string type = cfg.get_string("grid.type", "uniform");
        std::vector<Real> nodes;
        if (type == "uniform") nodes = RadialGrid::make_uniform(cfg.get_double("grid.rmax"), cfg.get_double("grid.dr"));
        else if (type == "regions" || type == "stretched")
            nodes = RadialGrid::make_regions(cfg.get_double("grid.rmax"), cfg.get_pairs("grid.regions"),
                                             cfg.get_double("grid.max_ratio", 1.05));
        else if (type == "file") nodes = RadialGrid::read_file(cfg.get_string("grid.file"));
        else throw std::runtime_error("grid: unknown grid.type '" + type + "' (uniform|regions|file)");
        grid.finalize(nodes, stream);
    }
    solver.init(grid, stream);
    const Real xu_max = cfg.get_double("xi.max", 0.0);
    const Real xu_min = cfg.get_double("xi.min");
    box.xi_min = xi_from_user(xu_max);
    box.dxi = cfg.get_double("xi.step");
    if (!(xu_max > xu_min) || !(box.dxi > 0)) throw std::runtime_error("xi: need xi.min < xi.max and xi.step > 0");
    box.nxi = static_cast<int>(std::lround((xu_max - xu_min) / box.dxi)) + 1;
    box.k0 = 0;
    box.nloc = box.nxi;
    dt = cfg.get_double("time.dt", 0.0);
    nsteps = cfg.get_int("time.steps", 0);
    t = cfg.get_double("time.start", 0.0);
    if (dt <= 0 && nsteps > 0) throw std::runtime_error("time.dt must be > 0");
    seed = cfg.get_int("plasma.seed", 1);
    const std::string shape = cfg.get_string("beams.xi_shape", "linear");
    if (shape == "ngp") box.ngp = 1;
    else if (shape == "linear") box.ngp = 0;
    else throw std::runtime_error("beams.xi_shape must be linear or ngp");
    auto names = cfg.get_list("plasma.species");
    if (names.empty()) names.push_back("electrons");
    bool has_positive = false;
    int nmobile = 0;
    for (const auto& nm : names) {
        species.push_back(std::make_unique<PlasmaSpecies>());
        species.back()->init(cfg, nm, grid, m1);
        if (species.back()->q > 0) has_positive = true;
        if (species.back()->mobile) ++nmobile;
    }
    if (nmobile > QZ_MAX_SPECIES) throw std::runtime_error("at most 2 mobile plasma species on this path");
    neutralize = cfg.get_bool("plasma.neutralize", !has_positive);
    for (const auto& nm : cfg.get_list("beams")) {
        beams.push_back(std::make_unique<Beam>());
        beams.back()->init(cfg, nm, m1, stream);
    }
    outdir = cfg.get_string("output.dir", "");
    out_every = cfg.get_int("output.every", 0);
    out_axis = cfg.get_bool("output.axis", true);
    (void)cfg.get_int("output.beam_every", 0);
    M = grid.N + 1;
    K = box.nloc;
    NCF = m1 ? 18 : 6;
    NCB = m1 ? 15 : 5;
    NCD = m1 ? 9 : 3;
    fld.alloc(static_cast<size_t>(K) * M * NCF);
    bsrc.alloc(static_cast<size_t>(K) * M * NCB);
    diag.alloc(static_cast<size_t>(K) * M * NCD);
    hip_check(hipMemsetAsync(fld.p, 0, sizeof(Real) * fld.n, stream), "memset");
    hip_check(hipMemsetAsync(diag.p, 0, sizeof(Real) * diag.n, stream), "memset");
    bg.alloc(6 * static_cast<size_t>(M));
    ws.alloc(qz_sweep_workspace_doubles(M));
    // workgroups of the sweep: one when the run fits one (at most 4096 plasma particles and solve sets that fit
    // one compute unit's LDS), else one per 4096 particles; sweep.groups (this library's key) forces a count
    {
        int np = 0;
        for (const auto& s : species)
            if (s->mobile) np += s->cap;
        const int want = cfg.get_int("sweep.groups", 0);
        if (want < 0) throw std::runtime_error("sweep.groups must be 0 (automatic) or a workgroup count");
        groups = want > 0 ? want : (qz_sweep_fits_one_group(m1, grid.N, np) ? 1 : (np + 4095) / 4096);
        if (groups > 1) mws.alloc(qz_sweep_multi_workspace_doubles(M, np, groups, m1 ? 1 : 0));
    }
    lost.alloc(QZ_MAX_SPECIES);
    counters.alloc(24);
    int maxnp = 0;
    for (const auto& b : beams) maxnp = std::max(maxnp, b->Np);
    dep_work.alloc(qz_beam_deposit_workspace_bytes(std::max(maxnp, 1), K, grid.N));
    step_n = cfg.get_int("time.first_step", 0);
    // keys that were never read: a key of this path that the setup does not need is reported, any other key
    // is an error before the first step (src/main.cpp:100-116)
    {
        const bool allow = cfg.get_bool("input.allow_unknown", false);
        std::string unknown;
        for (const auto& k : cfg.unused_keys()) {
            if (known_key(k)) std::fprintf(stderr, "WARNING: input key '%s' was not used\n", k.c_str());
            else unknown += " '" + k + "'";
        }
        if (!unknown.empty() && !allow)
            throw std::runtime_error("unknown input key(s) or keys outside this path:" + unknown + " (misspelt?)");
    }
    hip_check(hipStreamSynchronize(stream), "setup");
}

Simulation::~Simulation() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (aux) (void)hipStreamSynchronize(aux);
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : aux_ev)
        if (e) (void)hipEventDestroy(e);
    if (aux) (void)hipStreamDestroy(aux);
    if (stream) (void)hipStreamDestroy(stream);
}

// src/Simulation.cpp:636-795 (one rank, no sub-slicing)
void Simulation::step_fields() {
    prepare_sweep();
    hip_check(hipEventRecord(ev[0], stream), "event");
    launch_sweeps(d_args, &sweep_args, 1, stream);
    hip_check(hipEventRecord(ev[1], stream), "event");
    hip_check(hipEventSynchronize(ev[1]), "sync");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, ev[0], ev[1]), "elapsed");
    finish_sweep(ms);
}

// everything of step_fields before the slice loop (src/Simulation.cpp:636-726): beam sources, plasma load,
// background; fills sweep_args
void Simulation::prepare_sweep(bool timed) {
    // ---- beams -> (xi, r) arrays
    if (timed) hip_check(hipEventRecord(ev[4], stream), "event");
    qz_check(qz_zero_async(bsrc.p, sizeof(Real) * bsrc.n, stream));
    const qz_radial_grid g = grid.view();
    qz_check(qz_zero_async(counters.p, sizeof(unsigned long long) * counters.n, stream));
    for (const auto& b : beams) {
        const qz_beam_soa s = b->soa();
        qz_check(qz_beam_count_live(b->Np, b->a[6].p, counters.p + 20, stream));
        qz_check(qz_beam_deposit(&g, &box, m1 ? 1 : 0, b->q, b->Np, &s, bsrc.p, NCB, dep_work.p, stream));
    }
    if (timed) hip_check(hipEventRecord(ev[5], stream), "event");
    // ---- plasma at the head of the box, background
    const Real z_front = t - box.xi_min;
    unsigned long long sd = seed;
    for (auto& s : species) s->load(grid, z_front, sd++, stream);
    if (!bg_ready) {   // the loaded rings are the same every step, and so is the background
        std::vector<Real> bgt(3 * static_cast<size_t>(M), 0.0), bgr(3 * static_cast<size_t>(M), 0.0);
        for (auto& s : species)
            if (!s->mobile) s->deposit_rho(grid, bgt, bgr);
        if (neutralize) {
            std::vector<Real> a(3 * static_cast<size_t>(M), 0.0), b(3 * static_cast<size_t>(M), 0.0);
            for (auto& s : species)
                if (s->mobile) s->deposit_rho(grid, a, b);
            for (size_t j = 0; j < a.size(); ++j) { bgt[j] -= a[j]; bgr[j] -= b[j]; }
        }
        bg_host = bgt;   // a member: the asynchronous upload reads it after this function returns
        bg_host.insert(bg_host.end(), bgr.begin(), bgr.end());
        bg.upload(bg_host, stream);
        bg_ready = true;
    }
    qz_check(qz_zero_async(lost.p, sizeof(int) * lost.n, stream));
    // ---- the sweep
    qz_sweep_args A{};
    A.chi_out = chi_rec.p;   // null unless recording
    A.grid = g;
    A.modes = m1 ? 1 : 0;
    A.picard = picard;
    A.nslices = K;
    A.ngp = box.ngp;
    A.dxi = box.dxi;
    int nsp = 0;
    for (auto& s : species) {
        if (!s->mobile) continue;
        A.sp[nsp++] = s->state();
        A.ab_order = s->ab_order;
        for (int i = 0; i < 5; ++i) A.ab_c[i] = s->ab_c[i];
        A.delta_min = s->delta_min;
        A.max_qsa = s->max_qsa;
    }
    A.nspecies = nsp;
    if (nsp == 0) {   // no mobile species: nothing is pushed, the weights are not used
        make_ab(3, A.ab_c);
        A.ab_order = 3;
    }
    for (int op = 0; op < 4; ++op) {
        A.op_a[op] = solver.d_a[op].p; A.op_b[op] = solver.d_b[op].p; A.op_c[op] = solver.d_c[op].p;
    }
    A.flux_coef = solver.flux_coef;
    A.grad_m = solver.d_gm.p; A.grad_0 = solver.d_g0.p; A.grad_p = solver.d_gp.p; A.inv_r = solver.d_ir.p;
    A.fac_l0_m = solver.d_l0_m.p; A.fac_l0_cp = solver.d_l0_cp.p;
    A.fac_l1_m = solver.d_l1_m.p; A.fac_l1_cp = solver.d_l1_cp.p;
    A.bsrc = bsrc.p; A.ncb = NCB;
    A.bg = bg.p;
    A.fld = fld.p; A.ncf = NCF;
    A.diag = diag.p; A.ncd = NCD;
    A.lost = lost.p;
    A.ws = ws.p;
    A.stats = counters.p;
    A.groups = groups;
    A.mws = groups > 1 ? mws.p : nullptr;
    // cold plasma: the sweep loads the rings at the head itself
    A.load = nsp > 0 ? 1 : 0;
    nsp = 0;
    for (auto& s : species) {
        if (!s->mobile) continue;
        if (!s->device_rings()) A.load = 0;
        A.rings[nsp++] = {s->ppc, s->ntheta, s->nring, s->density, s->d_cs.p};
    }
    sweep_args = A;
}

// after the slice loop: removed-particle count, work counters, timings (ms_sweep measured by the caller)
void Simulation::finish_sweep(float sweep_ms, bool timed) {
    const int nsp = sweep_args.nspecies;
    sweep_launches = 1;
    check_multi_sweep(mws, groups, stream);
    std::vector<int> hl;
    lost.download(hl, stream);
    lost_last = 0;
    for (int i = 0; i < nsp; ++i) lost_last += hl[i];
    ms_sweep = sweep_ms;
    ms_beam_dep = 0;
    if (timed) hip_check(hipEventElapsedTime(&ms_beam_dep, ev[4], ev[5]), "elapsed");
    std::vector<unsigned long long> cnt;
    counters.download(cnt, stream);
    particle_slices = static_cast<double>(cnt[0]);
    for (int i = 0; i < QZ_NPHASE; ++i) phase_ticks[i] = static_cast<double>(cnt[1 + i]);
    beam_deposited = static_cast<double>(cnt[20]);
    swept = true;
}

// src/Simulation.cpp:1435-1456 (fixed dt)
void Simulation::push_beams() {
    const bool last = (step_n >= nsteps);
    ms_beam = 0;
    if (last) return;
    if (!have_prev) { dt_prev = dt; have_prev = true; }
    enqueue_push(true);
    hip_check(hipEventRecord(ev[3], stream), "event");
    hip_check(hipEventSynchronize(ev[3]), "sync");
    hip_check(hipEventElapsedTime(&ms_beam, ev[2], ev[3]), "elapsed");
    {
        unsigned long long v = 0;
        hip_check(hipMemcpyAsync(&v, counters.p + 21, sizeof(v), hipMemcpyDeviceToHost, stream), "counter");
        hip_check(hipStreamSynchronize(stream), "sync");
        beam_pushed = static_cast<double>(v);
    }
    t += dt;
    dt_prev = dt;
}

// device work of one beam push: live count, then the leapfrog start-up (first push only) and the kick-drift
void Simulation::enqueue_push(bool timed) {
    const Real dkick = Real(0.5) * (dt_prev + dt), ddrift = dt;
    const qz_radial_grid g = grid.view();
    qz_check(qz_zero_async(counters.p + 21, sizeof(unsigned long long), stream));
    for (auto& b : beams)
        if (!b->rigid) qz_check(qz_beam_count_live(b->Np, b->a[6].p, counters.p + 21, stream));
    if (timed) hip_check(hipEventRecord(ev[2], stream), "event");
    for (auto& b : beams) {
        if (b->rigid) continue;
        const qz_beam_soa s = b->soa();
        const Real qm = b->q / b->m;
        if (!b->started) {   // leapfrog start-up (src/Beam.cpp:900-905)
            qz_check(qz_beam_advance(&g, &box, m1 ? 1 : 0, qm, b->pusher, 0, b->Np, &s, fld.p, NCF, -0.5 * dkick,
                                     -0.5 * dkick, 0, stream));
            b->started = true;
        }
        qz_check(qz_beam_advance(&g, &box, m1 ? 1 : 0, qm, b->pusher, 0, b->Np, &s, fld.p, NCF, dkick, ddrift, 1, stream));
    }
}

// the device work of the next step equals that of the step before: background and plasma template made,
// every beam started (no leapfrog start-up), the same dt, a beam push to do and no output directory
bool Simulation::capturable() const {
    if (done || !bg_ready || !have_prev || dt_prev != dt || step_n >= nsteps) return false;
    if (!outdir.empty()) return false;   // the beam log (and field output) of every step is host work
    for (const auto& s : species)
        if (s->mobile && !s->have_template) return false;
    for (const auto& b : beams)
        if (!b->rigid && !b->started) return false;
    return true;
}

void Simulation::step() {
    if (done) throw std::runtime_error("the run has ended (time.steps reached)");
    step_fields();
    after_sweep();
}

// the rest of one iteration of Simulation::run after the sweep: output, beam push, step counter
void Simulation::after_sweep() {
    if (!outdir.empty()) write_beam_output();
    if (!outdir.empty() && out_every > 0 && step_n % out_every == 0) {
        char nm[32];
        std::snprintf(nm, sizeof(nm), "/fields_%06d.bin", step_n);
        write_fields(outdir + nm);
        if (out_axis) {
            std::snprintf(nm, sizeof(nm), "/axis_%06d.txt", step_n);
            write_axis(outdir + nm);
        }
    }
    const bool last = (step_n >= nsteps);
    push_beams();
    if (last) done = true;
    ++step_n;
}

// Simulation::write_beam_output for one rank (src/Simulation.cpp:1066-1078, header :259-261): every step, after
// the sweep and before the beam push, one line per beam with the moments of Beam::from_sums.
void Simulation::write_beam_output() {
    if (!beamlog.is_open()) {
        beamlog.open(outdir + "/beams.txt", std::ios::out | std::ios::trunc);
        if (!beamlog) throw std::runtime_error("cannot write " + outdir + "/beams.txt");
        beamlog << "# beam step t alive npart charge gamma_mean gamma_rms xi_mean xi_rms r_rms emit_nx"
                   " x_mean y_mean emit_ny sigma_x sigma_y\n";
    }
    for (const auto& b : beams) {
        const BeamDiag d = b->from_sums(b->sums_device(stream), t);
        beamlog << b->name << " " << step_n << " " << std::setprecision(10) << d.t << " " << d.alive << " " << d.npart
                << " " << d.charge << " " << d.gamma_mean << " " << d.gamma_rms << " " << d.xi_mean << " " << d.xi_rms
                << " " << d.r_rms << " " << d.emit_nx << " " << d.x_mean << " " << d.y_mean << " " << d.emit_ny << " "
                << d.sigma_x << " " << d.sigma_y << "\n";
    }
    beamlog.flush();
}

// ---------------------------------------------------------------------------- pipelined steps
// Steps of Simulation::run overlapped in time: step n sweeps slice range c as soon as step n-1 has pushed
// the beam particles that can reach it.  The result is the serial one bit for bit:
//  - the plasma is reloaded at the head every step and B+ restarts at zero there (src/Simulation.cpp:705-709),
//    so step n depends on step n-1 only through the beam particles;
//  - beam particles only slip to later slices (src/Beam.cpp:1357): the particles step n deposits into rows
//    <= r were pushed by step n-1 from rows <= r, and a range pushed once its fields exist, with every
//    particle marked (qz_beam_advance_slices), is pushed exactly as Beam::advance pushes it;
//  - the deposit of a row range is the full deposit restricted to those rows (node sums in particle-index
//    order), and step n overwrites rows of fld only after step n-1's pushes have read them.
// Every step in flight has its own stream, plasma arrays, workspace, bsrc and counters (a context); beam
// particles, fld and diag are shared.  Events order the ranges: deposit(n, c) waits for push(n-1, c+1), which
// pushes the particles that can land in rows up to cb[c+1] and follows the sweep of range c+1 of step n-1
// (so step n runs two ranges behind step n-1, and step n-1 has read the rows step n overwrites).
// The sweep launches read their argument blocks from one device array made before the first launch: a
// stream-ordered allocation per launch (as qz_sweep makes) could be handed from one stream to another while
// a kernel still reads it.
namespace {
struct StepCtx {
    hipStream_t st = nullptr;
    std::vector<std::array<DevBuf<Real>, 7>> sp;   // per mobile species: x y px py dl w hist
    DevBuf<Real> ws, bsrc;
    DevBuf<Real> mws;                               // runs of several workgroups: their workspace
    DevBuf<char> work;
    std::vector<hipEvent_t> pushed;                 // per slice range: this context's step pushed it
    ~StepCtx() {
        if (st) (void)hipStreamSynchronize(st);
        for (auto e : pushed)
            if (e) (void)hipEventDestroy(e);
        if (st) (void)hipStreamDestroy(st);
    }
};
}  // namespace

void Simulation::run_pipelined(int nrun, int nchunks, int depth, double& psl_total, double& pushed_total) {
    if (nrun < 1) throw std::runtime_error("run_pipelined: nrun < 1");
    if (!capturable())
        throw std::runtime_error("run_pipelined: the run must be started (every beam started, the same dt, no output "
                                 "directory, steps left)");
    if (step_n + nrun - 1 > nsteps) throw std::runtime_error("run_pipelined: more steps than time.steps leaves");
    // ranges of at least two slices: a range's deposit then needs only the previous step's pushes up to the
    // next range (a particle in the deposit window has its gather slice below the range after next)
    nchunks = std::max(1, std::min(nchunks, K / 2));
    depth = std::max(1, std::min(depth, nrun));
    if (groups > 1) {   // every step's launch of `groups` workgroups resident beside the others (workers_beside)
        const int ncu = device_cus();
        if (depth * ((groups + kXcds - 1) / kXcds) > ncu / kXcds)
            throw std::runtime_error("run_pipelined: " + std::to_string(depth) + " steps in flight of " +
                                     std::to_string(groups) + " workgroups each cannot all be resident; use at most " +
                                     std::to_string((ncu / kXcds) / ((groups + kXcds - 1) / kXcds)) + " in flight");
    }
    std::vector<int> cb(nchunks + 1);
    for (int c = 0; c <= nchunks; ++c) cb[c] = static_cast<int>((static_cast<long long>(K) * c) / nchunks);

    std::vector<PlasmaSpecies*> mob;
    for (auto& sp : species)
        if (sp->mobile) mob.push_back(sp.get());
    // the beams' windowed deposits go in one pass (qz_beam_deposit_beams); the window is at most one range + 1 rows
    std::vector<double> dq;
    std::vector<int> dnp;
    std::vector<qz_beam_soa> dsoa;
    std::vector<Beam*> dbeam;
    int dtot = 0;
    for (auto& b : beams)
        if (b->Np > 0) {
            dq.push_back(b->q);
            dnp.push_back(b->Np);
            dsoa.push_back(b->soa());
            dbeam.push_back(b.get());
            dtot += b->Np;
        }
    int wmax = 1;
    for (int c = 0; c < nchunks; ++c) wmax = std::max(wmax, cb[c + 1] - cb[c] + 1);
    const size_t dep_bytes = qz_beam_deposit_workspace_bytes(std::max(dtot, 1), std::min(wmax, K), grid.N);
    std::vector<std::unique_ptr<StepCtx>> ctx(depth);
    for (auto& cp : ctx) {
        cp.reset(new StepCtx());
        hip_check(hipStreamCreate(&cp->st), "stream");
        cp->sp.resize(mob.size());
        for (size_t s = 0; s < mob.size(); ++s) {
            for (int a = 0; a < 6; ++a) cp->sp[s][a].alloc(mob[s]->cap);
            cp->sp[s][6].alloc(mob[s]->hist.n);
        }
        cp->ws.alloc(ws.n);
        cp->bsrc.alloc(bsrc.n);
        if (groups > 1) {   // the barrier counter and the error flag start at zero
            cp->mws.alloc(mws.n);
            qz_check(qz_zero_async(cp->mws.p, 64 * sizeof(Real), stream));
        }
        cp->work.alloc(dep_bytes);
        cp->pushed.resize(nchunks);
        for (auto& e : cp->pushed) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
    }
    DevBuf<int> lostv(static_cast<size_t>(nrun) * QZ_MAX_SPECIES);
    DevBuf<unsigned long long> cntv(static_cast<size_t>(nrun) * 24);
    qz_check(qz_zero_async(lostv.p, sizeof(int) * lostv.n, stream));
    qz_check(qz_zero_async(cntv.p, sizeof(unsigned long long) * cntv.n, stream));
    for (auto& b : beams)
        if (b->pstep.n != static_cast<size_t>(b->Np)) {
            b->pstep.alloc(b->Np);
            qz_check(qz_zero_async(b->pstep.p, sizeof(int) * b->pstep.n, stream));
        }
    hip_check(hipStreamSynchronize(stream), "sync");   // the run's own stream is done before the contexts start
    std::vector<const int*> dpstep;   // the deposits leave out what their own step has pushed
    for (Beam* b : dbeam) dpstep.push_back(b->pstep.p);

    // the sweep arguments of every (step, range): the run's, with the step context's own arrays
    std::vector<qz_sweep_args> hargs(static_cast<size_t>(nrun) * nchunks);
    for (int i = 0; i < nrun; ++i) {
        StepCtx& C = *ctx[i % depth];
        qz_sweep_args A = sweep_args;
        A.bsrc = C.bsrc.p;
        A.ws = C.ws.p;
        A.mws = groups > 1 ? C.mws.p : nullptr;
        A.lost = lostv.p + static_cast<size_t>(i) * QZ_MAX_SPECIES;
        A.stats = cntv.p + static_cast<size_t>(i) * 24;
        for (size_t s = 0; s < mob.size(); ++s) {
            A.sp[s].x = C.sp[s][0].p; A.sp[s].y = C.sp[s][1].p; A.sp[s].px = C.sp[s][2].p;
            A.sp[s].py = C.sp[s][3].p; A.sp[s].dl = C.sp[s][4].p; A.sp[s].w = C.sp[s][5].p;
            A.sp[s].hist = C.sp[s][6].p;
        }
        for (int c = 0; c < nchunks; ++c) {
            A.k_begin = cb[c];
            A.k_end = cb[c + 1];
            if (qz_validate_sweep(A) != QZ_OK) throw std::runtime_error(qz_last_error());
            hargs[static_cast<size_t>(i) * nchunks + c] = A;
        }
    }
    DevBuf<qz_sweep_args> dargs(hargs.size());
    hip_check(hipMemcpy(dargs.p, hargs.data(), sizeof(qz_sweep_args) * hargs.size(), hipMemcpyHostToDevice), "arguments");
    // a range of a run of several workgroups: its block map (one run); the error flag is never cleared between
    // the ranges, so a barrier that gave up in any range of a step is reported below
    DevBuf<int> dmap;
    if (groups > 1) {
        const std::vector<int> map = qz_sweep_block_map(hargs.data(), 1);
        dmap.alloc(map.size());
        hip_check(hipMemcpy(dmap.p, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice), "block map");
    }

    const qz_radial_grid g = grid.view();
    const Real dkick = Real(0.5) * (dt_prev + dt), ddrift = dt;
    const int mm = m1 ? 1 : 0;
    const size_t row = static_cast<size_t>(M) * NCB;
    // HIP events around every range's sweep launch, on its stream: the kernel time of the timed steps
    struct LaunchEvents {
        std::vector<hipEvent_t> e;
        ~LaunchEvents() {
            for (auto x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } lev;
    lev.e.assign(4 * static_cast<size_t>(nrun) * nchunks, nullptr);   // sweep start, end; deposit start; push end
    for (auto& x : lev.e) hip_check(hipEventCreate(&x), "event");
    for (int i = 0; i < nrun; ++i) {
        StepCtx& C = *ctx[i % depth];
        StepCtx* P = i > 0 ? ctx[(i - 1) % depth].get() : nullptr;
        const int n = step_n + i;
        const bool last = n >= nsteps;   // the last sweep has no beam push (src/Simulation.cpp:1430-1446)
        const int marker = n + 1;
        hipStream_t st = C.st;
        // plasma at the head (the rings of the first load) with a fresh AB history; a cold plasma is loaded by
        // the first range's sweep itself
        for (size_t s = 0; s < mob.size() && !sweep_args.load; ++s) {
            const size_t bytes = sizeof(Real) * static_cast<size_t>(mob[s]->Np);
            const DevBuf<Real>* src[6] = {&mob[s]->t_x, &mob[s]->t_y, &mob[s]->t_px, &mob[s]->t_py, &mob[s]->t_dl,
                                          &mob[s]->t_w};
            for (int a = 0; a < 6; ++a)
                if (bytes) hip_check(hipMemcpyAsync(C.sp[s][a].p, src[a]->p, bytes, hipMemcpyDeviceToDevice, st), "plasma copy");
            qz_check(qz_zero_async(C.sp[s][6].p, sizeof(Real) * C.sp[s][6].n, st));
        }
        unsigned long long* cnt = cntv.p + static_cast<size_t>(i) * 24;
        // push range c: gather slices [cb[c] - 1, cb[c+1] - 1) (from 0 for the first, to K for the last), whose
        // stencil k, k+1 lies in slice ranges <= c; the range's last slice goes with the next push
        auto push_range = [&](int c) {
            const int lo = c == 0 ? 0 : cb[c] - 1, hi = c == nchunks - 1 ? K : cb[c + 1] - 1;
            for (auto& b : beams) {
                if (b->rigid || b->Np == 0) continue;
                const qz_beam_soa so = b->soa();
                qz_check(qz_beam_advance_slices(&g, &box, mm, b->q / b->m, b->pusher, b->Np, &so, fld.p, NCF, dkick, ddrift,
                                                1, lo, hi, marker, b->pstep.p, cnt + 21, st));
            }
            hip_check(hipEventRecord(C.pushed[c], st), "event");
        };
        for (int c = 0; c < nchunks; ++c) {
            const int c0 = cb[c], c1 = cb[c + 1];
            // beam sources of rows [d0, d1]: up to row c1, which the range's last slice differentiates
            if (P) hip_check(hipStreamWaitEvent(st, P->pushed[std::min(c + 1, nchunks - 1)], 0), "wait");
            const size_t ai = static_cast<size_t>(i) * nchunks + c;
            const size_t ne = 2 * static_cast<size_t>(nrun) * nchunks;   // the deposit and push marks follow
            hip_check(hipEventRecord(lev.e[ne + 2 * ai], st), "event");
            const int d0 = c == 0 ? 0 : c0 + 1, d1 = std::min(c1, K - 1);
            // the counts of Simulation::step (live beam particles before the deposit and before the push) are
            // taken by the kernels themselves: each window counts the live particles whose lower slice it holds,
            // each range's push the particles it marks
            if (d0 <= d1) {   // the window's rows are overwritten: no zero fill first
                qz_beam_grid bx = box;
                bx.k0 = box.k0 + d0;
                bx.nloc = d1 - d0 + 1;
                if (!dq.empty())
                    qz_check(qz_beam_deposit_window(&g, &bx, mm, static_cast<int>(dq.size()), dq.data(), dnp.data(),
                                                    dsoa.data(), C.bsrc.p + d0 * row, NCB, C.work.p, dpstep.data(),
                                                    marker, cnt + 20, st));
                else
                    qz_check(qz_zero_async(C.bsrc.p + d0 * row, sizeof(Real) * row * static_cast<size_t>(d1 - d0 + 1),
                                           st));
            }
            hip_check(hipEventRecord(lev.e[2 * ai], st), "event");
            if (groups > 1) {
                // steps in flight on their own streams: ordinary launches whose workgroups all fit the device.  The
                // barrier counter is zero: zeroed with the workspace, and put back to zero by each launch's last
                // workgroup out
                qz_check(qz_sweep_multi_launch(dargs.p + ai, dmap.p, hargs.data() + ai, 1, st, depth * groups));
            } else {
                qz_check(qz_sweep_launch(dargs.p + ai, hargs.data() + ai, 1, st));
            }
            hip_check(hipEventRecord(lev.e[2 * ai + 1], st), "event");
            if (!last) push_range(c);
            hip_check(hipEventRecord(lev.e[ne + 2 * ai + 1], st), "event");
        }
    }
    for (auto& cp : ctx) hip_check(hipStreamSynchronize(cp->st), "sync");
    if (groups > 1)
        for (auto& cp : ctx) check_multi_sweep(cp->mws, groups, cp->st);

    // host side of the steps: totals, the last step's counters, the run's arrays as after the last step
    std::vector<unsigned long long> hc;
    cntv.download(hc, stream);
    std::vector<int> hl;
    lostv.download(hl, stream);
    psl_total = pushed_total = 0;
    for (int i = 0; i < nrun; ++i) {
        psl_total += static_cast<double>(hc[static_cast<size_t>(i) * 24]);
        if (step_n + i < nsteps) pushed_total += static_cast<double>(hc[static_cast<size_t>(i) * 24 + 21]);
    }
    const size_t li = static_cast<size_t>(nrun - 1);
    particle_slices = static_cast<double>(hc[li * 24]);
    for (int k = 0; k < QZ_NPHASE; ++k) phase_ticks[k] = static_cast<double>(hc[li * 24 + 1 + k]);
    beam_deposited = static_cast<double>(hc[li * 24 + 20]);
    beam_pushed = static_cast<double>(hc[li * 24 + 21]);
    lost_last = 0;
    for (size_t s = 0; s < mob.size(); ++s) lost_last += hl[li * QZ_MAX_SPECIES + s];
    StepCtx& Lc = *ctx[(nrun - 1) % depth];
    for (size_t s = 0; s < mob.size(); ++s) {
        DevBuf<Real>* dst[7] = {&mob[s]->x, &mob[s]->y, &mob[s]->px, &mob[s]->py, &mob[s]->dl, &mob[s]->w, &mob[s]->hist};
        for (int a = 0; a < 7; ++a)
            hip_check(hipMemcpyAsync(dst[a]->p, Lc.sp[s][a].p, sizeof(Real) * dst[a]->n, hipMemcpyDeviceToDevice, stream),
                      "plasma copy");
    }
    hip_check(hipMemcpyAsync(bsrc.p, Lc.bsrc.p, sizeof(Real) * bsrc.n, hipMemcpyDeviceToDevice, stream), "bsrc copy");
    hip_check(hipStreamSynchronize(stream), "sync");
    double launch_ms = 0;
    const size_t nl = static_cast<size_t>(nrun) * nchunks;
    pipe_timeline.assign(4 * nl, 0.0);
    for (size_t ai = 0; ai < nl; ++ai) {
        float ms = 0, t[4] = {0, 0, 0, 0};
        hip_check(hipEventElapsedTime(&ms, lev.e[2 * ai], lev.e[2 * ai + 1]), "elapsed");
        const hipEvent_t marks[4] = {lev.e[2 * nl + 2 * ai], lev.e[2 * ai], lev.e[2 * ai + 1], lev.e[2 * nl + 2 * ai + 1]};
        for (int m = 0; m < 4; ++m) hip_check(hipEventElapsedTime(&t[m], lev.e[2 * nl], marks[m]), "elapsed");
        for (int m = 0; m < 4; ++m) pipe_timeline[4 * ai + m] = t[m];
        launch_ms += ms;
    }
    pipe_chunks = nchunks;
    ms_sweep = static_cast<float>(launch_ms / nrun);   // per step: the sum of its range launches' kernel times
    sweep_launches = nchunks;
    ms_beam = ms_beam_dep = 0;
    swept = true;
    for (int i = 0; i < nrun; ++i) {   // Simulation::push_beams and after_sweep, per step
        if (step_n < nsteps) {
            t += dt;
            dt_prev = dt;
        } else {
            done = true;
        }
        ++step_n;
    }
}

// fields_NNNNNN.bin version 3 (src/Simulation.cpp:873-993), all components, one rank
// fields_*.bin, format v3 (src/Simulation.cpp:873-993): header, r[M], xi[K] increasing (head last), then per
// component its name[16] and [K][M] doubles; with m = 1 the _c / _s parts are 2 Re / -2 Im of the mode (vectors:
// e_r and e_theta from the +1/-1 helical parts, :902-928).  Streamed: each component is formed in HBM from fld,
// diag or bsrc and copied to the host alone, so the host holds K x M doubles instead of the whole field store.
// ---------------------------------------------------------------------------- the adjoint of one step (m = 0)
namespace {
void grow(DevBuf<Real>& b, size_t n) {
    if (b.n < n) b.alloc(n);
}
}  // namespace

Simulation::AdjointBufs::~AdjointBufs() {
    if (st2) (void)hipStreamSynchronize(st2);
    for (auto e : ev)
        if (e) (void)hipEventDestroy(e);
    if (st2) (void)hipStreamDestroy(st2);
}

void Simulation::spread_value_grad(int bi, int every, double& J, std::vector<double>& grad) {
    if (done) throw std::runtime_error("the run has ended (time.steps reached)");
    if (m1) throw std::runtime_error("adjoint: modes = 0 only");
    if (groups > 1) throw std::runtime_error("adjoint: a one-workgroup sweep only (sweep.groups = 1)");
    if (bi < 0 || bi >= static_cast<int>(beams.size())) throw std::runtime_error("adjoint: no such beam");
    if (every < 1) throw std::runtime_error("adjoint: checkpoint interval must be at least one slice");
    Beam& B = *beams[bi];
    if (B.pwl.empty()) throw std::runtime_error("adjoint: " + B.name + " has no weight profile (pwl_weights)");
    if (B.rigid) throw std::runtime_error("adjoint: " + B.name + " is rigid");
    if (B.pusher != 0) throw std::runtime_error("adjoint: the Vay beam pusher only");
    if (step_n >= nsteps) throw std::runtime_error("adjoint: the step must push the beams (time.steps)");
    if (!adj) adj = std::make_unique<AdjointBufs>();
    AdjointBufs& W = *adj;
    const size_t Nb = static_cast<size_t>(B.Np);
    // the beam before its push, and the kick the push will give
    for (int v = 0; v < 7; ++v) {
        grow(W.before[v], std::max<size_t>(Nb, 1));
        if (Nb) hip_check(hipMemcpyAsync(W.before[v].p, B.a[v].p, sizeof(Real) * Nb, hipMemcpyDeviceToDevice, stream),
                          "beam copy");
    }
    const bool started = B.started;
    const double dkick = have_prev ? 0.5 * (dt_prev + dt) : dt;
    // ---- the step, its sweep in checkpointed segments
    prepare_sweep();
    const qz_sweep_args A = sweep_args;
    const int nb = static_cast<int>(beams.size());
    const int nseg = (K + every - 1) / every;
    const size_t msg = qz_message_doubles(&A, nb);
    grow(W.ckpt, msg * static_cast<size_t>(nseg));
    hip_check(hipEventRecord(ev[0], stream), "event");
    for (int s = 0; s < nseg; ++s) {
        const int lo = s * every, hi = std::min(K, lo + every);
        // segment 0 of a cold plasma loads its rings in the kernel: no state to keep
        if (s > 0 || !A.load) qz_check(qz_pack_message(&A, nb, step_n, dt, 0, lo, W.ckpt.p + msg * s, stream));
        qz_sweep_args a = A;
        a.k_begin = lo;
        a.k_end = hi;
        a.load = lo == 0 ? A.load : 0;
        launch_sweeps(d_args, &a, 1, stream);
    }
    hip_check(hipEventRecord(ev[1], stream), "event");
    hip_check(hipEventSynchronize(ev[1]), "sync");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, ev[0], ev[1]), "elapsed");
    finish_sweep(ms);
    sweep_launches = nseg;
    ms_adj_forward = ms;
    after_sweep();
    // ---- the objective
    const std::array<double, 17> S = B.sums_device(stream);
    const BeamDiag d = B.from_sums(S, t);
    if (!(d.gamma_mean > 0) || !(d.gamma_rms > 0)) throw std::runtime_error("adjoint: the beam has no energy spread");
    J = d.gamma_rms / d.gamma_mean;
    hip_check(hipEventRecord(ev[0], stream), "event");
    grow(W.lam_gamma, std::max<size_t>(Nb, 1));
    grow(W.lam_w, std::max<size_t>(Nb, 1));
    const qz_beam_soa after = B.soa();
    qz_check(qz_beam_spread_adjoint(B.Np, &after, d.gamma_mean, d.gamma_rms, S[0], W.lam_gamma.p, W.lam_w.p, stream));
    // ---- the kick
    const qz_radial_grid g = grid.view();
    const qz_beam_soa before{W.before[0].p, W.before[1].p, W.before[2].p, W.before[3].p, W.before[4].p,
                             W.before[5].p, W.before[6].p};
    const size_t nf = static_cast<size_t>(K) * M * NCF, nbs = static_cast<size_t>(K) * M * NCB;
    grow(W.lam_fld, nf);
    grow(W.lam_bsrc, nbs);
    qz_check(qz_zero_async(W.lam_fld.p, sizeof(Real) * nf, stream));
    qz_check(qz_zero_async(W.lam_bsrc.p, sizeof(Real) * nbs, stream));
    qz_check(qz_beam_kick_adjoint(&g, &box, B.q / B.m, B.pusher, B.Np, &before, fld.p, NCF, dkick, started ? 1 : 0,
                                  W.lam_gamma.p, W.lam_fld.p, stream));
    // ---- the reverse sweep, segment by segment from the box's tail.  Segment s's reverse sweep (stream st2)
    // overlaps the recomputation of segment s - 1 (the run's stream): they touch different slices and buffers.
    int np = 0;
    for (int s = 0; s < A.nspecies; ++s) np += A.sp[s].np;
    const size_t nring = static_cast<size_t>(A.ab_order) * 5 * np;
    grow(W.ring, std::max<size_t>(nring, 1));
    for (auto& tb : W.traj) grow(tb, std::max<size_t>(static_cast<size_t>(every) * 6 * np, 1));
    grow(W.work, qz_sweep_adjoint_workspace_doubles(M, np));
    qz_check(qz_zero_async(W.ring.p, sizeof(Real) * std::max<size_t>(nring, 1), stream));
    if (W.ticks.n < QZ_ADJ_NPHASE) W.ticks.alloc(QZ_ADJ_NPHASE);
    qz_check(qz_zero_async(W.ticks.p, sizeof(unsigned long long) * QZ_ADJ_NPHASE, stream));
    if (!W.st2) hip_check(hipStreamCreateWithFlags(&W.st2, hipStreamNonBlocking), "stream");
    while (W.ev.size() < 2 * static_cast<size_t>(nseg)) {
        hipEvent_t e = nullptr;
        hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
        W.ev.push_back(e);
    }
    auto recompute = [&](int s) {
        const int lo = s * every, hi = std::min(K, lo + every);
        qz_sweep_args a = A;
        if (s > 0 || !A.load) qz_check(qz_unpack_message(&a, nb, W.ckpt.p + msg * s, nullptr, stream));
        a.k_begin = lo;
        a.k_end = hi;
        a.load = (s == 0 && A.load) ? 1 : 0;
        a.traj = W.traj[s & 1].p;
        a.stats = nullptr;
        launch_sweeps(d_args, &a, 1, stream);
        hip_check(hipEventRecord(W.ev[2 * s], stream), "event");
        return a;
    };
    qz_sweep_args a = recompute(nseg - 1);
    for (int s = nseg - 1; s >= 0; --s) {
        const int lo = s * every, hi = std::min(K, lo + every);
        hip_check(hipStreamWaitEvent(W.st2, W.ev[2 * s], 0), "stream wait");
        const qz_adjoint_args J{W.traj[s & 1].p, lo, hi, W.lam_fld.p, W.lam_bsrc.p, W.ring.p, W.work.p, W.ticks.p};
        qz_check(qz_sweep_adjoint(&a, &J, W.st2));
        hip_check(hipEventRecord(W.ev[2 * s + 1], W.st2), "event");
        if (s > 0) {
            // segment s - 1 reuses the trajectory buffer of segment s + 1: after that segment's reverse sweep
            if (s + 1 < nseg) hip_check(hipStreamWaitEvent(stream, W.ev[2 * (s + 1) + 1], 0), "stream wait");
            a = recompute(s - 1);
        }
    }
    hip_check(hipStreamWaitEvent(stream, W.ev[1], 0), "stream wait");
    // ---- the beam deposit and the profile
    qz_check(qz_beam_deposit_adjoint(&g, &box, B.q, B.Np, &before, W.lam_bsrc.p, NCB, W.lam_w.p, stream));
    if (W.u_beam != bi || W.u.n < Nb) {
        std::vector<Real> u(B.xi_init.size());
        for (size_t i = 0; i < u.size(); ++i) u[i] = xi_to_user(B.xi_init[i]);
        W.u.upload(u, stream);
        W.u_beam = bi;
    }
    const int S_ = static_cast<int>(B.pwl.size());
    grow(W.grad, static_cast<size_t>(S_));
    qz_check(qz_profile_gradient(B.Np, W.u.p, W.lam_w.p, S_, B.pwl_u0, B.pwl_u1, B.pwl_wmac, W.grad.p, stream));
    hip_check(hipEventRecord(ev[1], stream), "event");
    grad.assign(static_cast<size_t>(S_), 0.0);
    hip_check(hipMemcpyAsync(grad.data(), W.grad.p, sizeof(Real) * S_, hipMemcpyDeviceToHost, stream), "gradient");
    hip_check(hipStreamSynchronize(stream), "sync");
    hip_check(hipEventElapsedTime(&ms_adj_reverse, ev[0], ev[1]), "elapsed");
    std::vector<unsigned long long> tk;
    W.ticks.download(tk, stream);
    for (int i = 0; i < QZ_ADJ_NPHASE; ++i) W.last_ticks[i] = static_cast<double>(tk[i]);
}

void Simulation::write_fields(const std::string& fn) {
    struct Comp { const char* name; int src; int c0; double s0; int c1; double s1; };   // src: 0 fld, 1 diag, 2 bsrc
    std::vector<Comp> comps = {
        {"psi", 1, D_PSI0, 1, 0, 0}, {"ez", 0, F_EZ0, 1, 0, 0}, {"er", 0, F_EP1R, 1, 0, 0}, {"eth", 0, F_EP1I, 1, 0, 0},
        {"br", 0, F_BP1R, 1, 0, 0}, {"bth", 0, F_BP1I, 1, 0, 0}, {"bz", 0, F_BZ0, 1, 0, 0}, {"ne", 1, D_NE0, 1, 0, 0},
        {"ni", 1, D_NI0, 1, 0, 0}, {"rhob", 2, B_RHO0, 1, 0, 0},
    };
    if (m1) {
        const std::vector<Comp> more = {
            {"psi_c", 1, D_PSI1R, 2, 0, 0}, {"psi_s", 1, D_PSI1I, -2, 0, 0},
            {"ez_c", 0, F_EZ1R, 2, 0, 0}, {"ez_s", 0, F_EZ1I, -2, 0, 0},
            {"bz_c", 0, F_BZ1R, 2, 0, 0}, {"bz_s", 0, F_BZ1I, -2, 0, 0},
            {"ne_c", 1, D_NE1R, 2, 0, 0}, {"ne_s", 1, D_NE1I, -2, 0, 0},
            {"ni_c", 1, D_NI1R, 2, 0, 0}, {"ni_s", 1, D_NI1I, -2, 0, 0},
            {"rhob_c", 2, B_RHO1R, 2, 0, 0}, {"rhob_s", 2, B_RHO1I, -2, 0, 0},
            {"er_c", 0, F_EP0R, 1, F_EP2R, 1}, {"er_s", 0, F_EP0I, 1, F_EP2I, -1},
            {"eth_c", 0, F_EP0I, 1, F_EP2I, 1}, {"eth_s", 0, F_EP2R, 1, F_EP0R, -1},
            {"br_c", 0, F_BP0R, 1, F_BP2R, 1}, {"br_s", 0, F_BP0I, 1, F_BP2I, -1},
            {"bth_c", 0, F_BP0I, 1, F_BP2I, 1}, {"bth_s", 0, F_BP2R, 1, F_BP0R, -1},
        };
        comps.insert(comps.end(), more.begin(), more.end());
    }
    std::FILE* fp = std::fopen(fn.c_str(), "wb");
    if (!fp) throw std::runtime_error("cannot write " + fn);
    struct Closer { std::FILE* f; ~Closer() { std::fclose(f); } } closer{fp};
    const int32_t hdr[4] = {3, M, K, m1 ? 1 : 0};
    std::fwrite(hdr, 4, 4, fp);
    std::fwrite(&t, 8, 1, fp);
    std::fwrite(grid.r.data(), 8, M, fp);
    for (int k = 0; k < K; ++k) {
        const double xx = xi_to_user(box.xi_min + (K - 1 - k) * box.dxi);
        std::fwrite(&xx, 8, 1, fp);
    }
    const int32_t nc = static_cast<int32_t>(comps.size());
    std::fwrite(&nc, 4, 1, fp);
    DevBuf<Real> dcomp(static_cast<size_t>(K) * M);
    std::vector<Real> hcomp;
    const Real* srcs[3] = {fld.p, diag.p, bsrc.p};
    const int ncs[3] = {NCF, NCD, NCB};
    for (const Comp& c : comps) {
        char name[16];
        std::memset(name, 0, sizeof(name));
        std::strncpy(name, c.name, 15);
        std::fwrite(name, 1, 16, fp);
        qz_check(qz_field_component(srcs[c.src], ncs[c.src], K, M, c.c0, c.s0, c.c1, c.s1, dcomp.p, stream));
        dcomp.download(hcomp, stream);
        std::fwrite(hcomp.data(), 8, hcomp.size(), fp);
    }
}

// axis_*.txt (src/Simulation.cpp:995-1029, one rank): a 256-byte header of two padded lines, then one fixed-width
// line per slice from the tail to the head: xi, Ez(0), psi(0), ne(0), rho_beam(0), (Er - Bth)/r at the first
// node off the axis, and with m = 1 the axis values of W = E_perp + z x B_perp.  Only the first two nodes of
// every slice are copied from HBM (strided copies), not the field store.
void Simulation::write_axis(const std::string& fn) {
    std::vector<Real> F(static_cast<size_t>(K) * 2 * NCF), D(static_cast<size_t>(K) * NCD), B(static_cast<size_t>(K) * NCB);
    hip_check(hipMemcpy2DAsync(F.data(), 16 * NCF, fld.p, 8 * static_cast<size_t>(M) * NCF, 16 * NCF, K,
                               hipMemcpyDeviceToHost, stream), "axis copy");
    hip_check(hipMemcpy2DAsync(D.data(), 8 * NCD, diag.p, 8 * static_cast<size_t>(M) * NCD, 8 * NCD, K,
                               hipMemcpyDeviceToHost, stream), "axis copy");
    hip_check(hipMemcpy2DAsync(B.data(), 8 * NCB, bsrc.p, 8 * static_cast<size_t>(M) * NCB, 8 * NCB, K,
                               hipMemcpyDeviceToHost, stream), "axis copy");
    hip_check(hipStreamSynchronize(stream), "sync");
    auto f = [&](int k, int j, int c) { return F[(static_cast<size_t>(k) * 2 + j) * NCF + c]; };
    auto d = [&](int k, int c) { return D[static_cast<size_t>(k) * NCD + c]; };
    auto b = [&](int k, int c) { return B[static_cast<size_t>(k) * NCB + c]; };
    const double r1 = grid.r[1];
    const int ncol = m1 ? 8 : 6;
    std::string h = "# t = " + std::to_string(t);
    h.resize(127, ' ');
    h += '\n';
    std::string h2 = "# xi  Ez(0)  psi(0)  ne(0)  rho_beam(0)  (Er-Bth)/r|r1";
    if (m1) h2 += "  Wx(0)  Wy(0)  [W = E_perp + z x B_perp on the axis, from m=1]";
    h2.resize(127, ' ');
    h2 += '\n';
    h += h2;
    std::string buf;
    buf.reserve(static_cast<size_t>(18) * ncol * K);
    char cell[32];
    for (int k = K - 1; k >= 0; --k) {   // increasing user xi: the head of the box is the last line
        double v[8] = {xi_to_user(box.xi_min + k * box.dxi), f(k, 0, F_EZ0), d(k, D_PSI0), d(k, D_NE0), b(k, B_RHO0),
                       (f(k, 1, F_EP1R) - f(k, 1, F_BP1I)) / r1, 0, 0};
        if (m1) { v[6] = f(k, 0, F_EP0R) - f(k, 0, F_BP0I); v[7] = f(k, 0, F_EP0I) + f(k, 0, F_BP0R); }
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

void Simulation::snapshot() {
    auto sn = std::make_unique<Snap>();
    sn->beams.resize(beams.size());
    for (size_t b = 0; b < beams.size(); ++b) {
        for (int c = 0; c < 7; ++c) {
            const DevBuf<Real>& a = beams[b]->a[c];
            sn->beams[b][c].alloc(a.n);
            if (a.n) hip_check(hipMemcpyAsync(sn->beams[b][c].p, a.p, sizeof(Real) * a.n, hipMemcpyDeviceToDevice, stream), "snapshot");
        }
        sn->started.push_back(beams[b]->started ? 1 : 0);
        sn->refs.push_back({beams[b]->gref, beams[b]->xiref});
        sn->pwl.push_back(beams[b]->pwl);
    }
    sn->t = t; sn->step_n = step_n; sn->done = done; sn->have_prev = have_prev; sn->swept = swept; sn->dt_prev = dt_prev;
    hip_check(hipStreamSynchronize(stream), "sync");
    snap = std::move(sn);
}

void Simulation::rewind() {
    if (!snap) throw std::runtime_error("rewind: no snapshot");
    for (size_t b = 0; b < beams.size(); ++b) {
        Beam& B = *beams[b];
        for (int c = 0; c < 7; ++c) {
            const DevBuf<Real>& a = snap->beams[b][c];
            if (a.n != B.a[c].n) throw std::runtime_error("rewind: beam arrays changed size");
            if (a.n) hip_check(hipMemcpyAsync(B.a[c].p, a.p, sizeof(Real) * a.n, hipMemcpyDeviceToDevice, stream), "rewind");
        }
        B.started = snap->started[b] != 0;
        B.gref = snap->refs[b][0];
        B.xiref = snap->refs[b][1];
        B.pwl = snap->pwl[b];
        B.pstep.release();   // the pipelined step markers restart with the step count
    }
    t = snap->t; step_n = snap->step_n; done = snap->done; have_prev = snap->have_prev; swept = snap->swept;
    dt_prev = snap->dt_prev;
    hip_check(hipStreamSynchronize(stream), "sync");
}

}  // namespace qz

// ============================================================================ qz_sim_* C-ABI
struct qz_sim {
    std::unique_ptr<qz::Simulation> s;
};

#define QZ_SIM_TRY(...)                                     \
    try {                                                   \
        __VA_ARGS__;                                        \
        return QZ_OK;                                       \
    } catch (const std::exception& e) {                     \
        return qz_fail(QZ_EINVAL, e.what());                \
    }

int qz_fail(int code, const std::string& msg);

// a null run handle (or a null output pointer) is a caller error, reported like any other
#define QZ_NEED(cond, fn)                                                        \
    do {                                                                         \
        if (!(cond)) return qz_fail(QZ_EINVAL, std::string(fn) + ": null argument"); \
    } while (0)

// The deck reader alone (host only, no device work): every key of the deck after the overrides, one
// "key = value" line each in key order.  Returns the bytes of the text (without the terminating 0), or a
// negative code; the text is written when cap exceeds that length.
extern "C" long qz_deck_keys(const char* deck, const char* overrides, char* out, long cap) {
    if (!deck) return qz_fail(QZ_EINVAL, "qz_deck_keys: null deck");
    try {
        qz::Config cfg;
        cfg.parse(deck);
        cfg.overrides(overrides ? overrides : "");
        const std::string text = cfg.dump();
        const long n = static_cast<long>(text.size());
        if (out && cap > n) std::memcpy(out, text.c_str(), static_cast<size_t>(n) + 1);
        return n;
    } catch (const std::exception& e) {
        return qz_fail(QZ_EINVAL, e.what());
    }
}

extern "C" int qz_sim_create(const char* deck, const char* overrides, qz_sim** out) {
    if (!deck || !out) return qz_fail(QZ_EINVAL, "qz_sim_create: null argument");
    QZ_SIM_TRY({
        auto* h = new qz_sim;
        try {
            h->s = std::make_unique<qz::Simulation>(deck, overrides ? overrides : "");
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    })
}
extern "C" void qz_sim_destroy(qz_sim* s) { delete s; }
extern "C" int qz_sim_info(const qz_sim* h, int* info) {
    if (!h || !info) return qz_fail(QZ_EINVAL, "qz_sim_info: null argument");
    const auto& s = *h->s;
    info[0] = s.M; info[1] = s.K; info[2] = s.m1 ? 1 : 0; info[3] = s.NCF; info[4] = s.NCB; info[5] = s.NCD;
    info[6] = static_cast<int>(s.species.size()); info[7] = static_cast<int>(s.beams.size());
    info[8] = s.nsteps; info[9] = s.step_n;
    return QZ_OK;
}
extern "C" int qz_sim_step(qz_sim* h) {
    QZ_NEED(h, "qz_sim_step");
    QZ_SIM_TRY(h->s->step())
}

// Ensures the run's second launch stream and its fork/join events exist.
static void ensure_aux(qz::Simulation& s) {
    if (s.aux) return;
    qz::hip_check(hipStreamCreate(&s.aux), "stream");
    for (auto& e : s.aux_ev) qz::hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
}

// One step of n independent runs with their sweeps launched together: the one-workgroup runs in one persistent
// launch whose workgroups take them from a queue, longest first, and the multi-workgroup runs (groups > 1) in
// one launch beside it (qz::MixedSweep).  Every run's preparation is enqueued on its own stream; the launches
// wait for all of them on the first run's stream; every run then finishes and pushes its beams on its own
// stream.
static int step_batched(qz_sim** hs, int n, double* batch_ms, int max_workers) {
    if (!hs || n <= 0) return qz_fail(QZ_EINVAL, "qz_sim_step_batched: no runs");
    QZ_SIM_TRY({
        std::vector<qz::Simulation*> runs(n);
        for (int i = 0; i < n; ++i) {
            if (!hs[i]) throw std::runtime_error("qz_sim_step_batched: null run");
            if (hs[i]->s->done) throw std::runtime_error("qz_sim_step_batched: a run has ended (time.steps reached)");
            if (hs[i]->s->m1 != hs[0]->s->m1) throw std::runtime_error("qz_sim_step_batched: runs differ in modes");
            runs[i] = hs[i]->s.get();
        }
        qz::Simulation& s0 = *hs[0]->s;
        for (int i = 0; i < n; ++i) {
            qz::Simulation& s = *hs[i]->s;
            s.prepare_sweep();
            qz::hip_check(hipEventRecord(s.ev[0], s.stream), "event");
            if (i > 0) qz::hip_check(hipStreamWaitEvent(s0.stream, s.ev[0], 0), "stream wait");
        }
        const qz::MixedSweep P = qz::plan_mixed(runs.data(), n, max_workers);
        if (!P.multi.empty() && !P.fused) ensure_aux(s0);
        struct Events {   // released on every exit path
            hipEvent_t e[2] = {nullptr, nullptr};
            ~Events() { for (auto x : e) if (x) (void)hipEventDestroy(x); }
        } ev;
        qz::hip_check(hipEventCreate(&ev.e[0]), "event");
        qz::hip_check(hipEventCreate(&ev.e[1]), "event");
        hipEvent_t e0 = ev.e[0], e1 = ev.e[1];
        qz::upload_mixed(P, s0.d_args, s0.stream, false);
        qz::hip_check(hipEventRecord(e0, s0.stream), "event");
        if (P.multi.empty() || P.fused) {
            qz::enqueue_mixed(P, s0.d_args.p, s0.stream, s0.aux, s0.aux_ev[0], s0.aux_ev[1]);
        } else {
            // two streams of one process can share a hardware queue, and then run one launch after the other;
            // a graph's independent branches run on streams of their own, so the two launches go through a
            // one-off graph, as in a captured step
            struct OneOff {
                hipGraph_t g = nullptr;
                hipGraphExec_t x = nullptr;
                ~OneOff() {
                    if (x) (void)hipGraphExecDestroy(x);
                    if (g) (void)hipGraphDestroy(g);
                }
            } oo;
            qz::hip_check(hipStreamBeginCapture(s0.stream, hipStreamCaptureModeThreadLocal), "begin capture");
            try {
                qz::enqueue_mixed(P, s0.d_args.p, s0.stream, s0.aux, s0.aux_ev[0], s0.aux_ev[1]);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(s0.stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            qz::hip_check(hipStreamEndCapture(s0.stream, &oo.g), "end capture");
            qz::hip_check(hipGraphInstantiate(&oo.x, oo.g, nullptr, nullptr, 0), "graph instantiate");
            qz::hip_check(hipGraphLaunch(oo.x, s0.stream), "graph launch");
            qz::hip_check(hipEventRecord(e1, s0.stream), "event");
            qz::hip_check(hipEventSynchronize(e1), "sync");   // before the one-off graph is destroyed
        }
        qz::hip_check(hipEventRecord(e1, s0.stream), "event");
        qz::hip_check(hipEventSynchronize(e1), "sync");
        float ms = 0;   // every multi-workgroup run's error flag is checked by its finish_sweep
        qz::hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
        if (batch_ms) *batch_ms = ms;
        // the rest of every run's step (counters, output, beam push) touches only that run's stream and
        // buffers: a few host threads do the runs side by side
        std::vector<std::exception_ptr> errs(static_cast<size_t>(n));
        std::atomic<int> next{0};
        auto work = [&]() {
            for (int i = next++; i < n; i = next++) {
                try {
                    qz::Simulation& s = *hs[i]->s;
                    s.finish_sweep(ms);
                    s.after_sweep();
                } catch (...) {
                    errs[static_cast<size_t>(i)] = std::current_exception();
                }
            }
        };
        const int nth = std::min(n, 8);
        std::vector<std::thread> pool;
        for (int t = 1; t < nth; ++t) pool.emplace_back(work);
        work();
        for (auto& th : pool) th.join();
        for (auto& e : errs)
            if (e) std::rethrow_exception(e);
    })
}

extern "C" int qz_sim_step_batched(qz_sim** hs, int n, double* batch_ms) { return step_batched(hs, n, batch_ms, 0); }

// A fixed set of runs stepped together.  Every step launches all sweeps together (qz_sim_step_batched); once a
// step repeats the device work of the one before in every run (Simulation::capturable), the whole step (each
// run's beam deposit and plasma reload, the sweep launches, each run's beam push) is captured once as a HIP
// graph on the batch's stream and replayed.  Not in the reference.
struct qz_batch {
    static constexpr int kBranches = 8;   // streams the per-run work of a captured step is spread over
    std::vector<qz_sim*> sims;
    bool use_graph = true;
    int max_workers = 0;                       // cap on the one-workgroup launch's workgroups (0: none)
    hipStream_t st = nullptr, aux = nullptr;   // aux: the multi-workgroup launch of a captured step
    hipStream_t br[kBranches] = {};
    hipEvent_t fork[2] = {}, join[2][kBranches] = {}, mfork = nullptr, mjoin = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    qz::MixedSweep plan;                     // the captured step's launches
    qz::DevBuf<qz_sweep_args> dargs;         // their device arguments, read by the graph's launches
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    long replays = 0;
    void drop_graph() {   // the next replayable step captures again
        if (st) (void)hipStreamSynchronize(st);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~qz_batch() {
        drop_graph();
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (mfork) (void)hipEventDestroy(mfork);
        if (mjoin) (void)hipEventDestroy(mjoin);
        for (int f = 0; f < 2; ++f) {
            if (fork[f]) (void)hipEventDestroy(fork[f]);
            for (auto e : join[f])
                if (e) (void)hipEventDestroy(e);
        }
        for (auto b : br)
            if (b) (void)hipStreamDestroy(b);
        if (aux) (void)hipStreamDestroy(aux);
        if (st) (void)hipStreamDestroy(st);
    }
};

extern "C" int qz_batch_create(qz_sim** hs, int n, int use_graph, qz_batch** out) {
    if (!hs || n <= 0 || !out) return qz_fail(QZ_EINVAL, "qz_batch_create: no runs");
    QZ_SIM_TRY({
        std::unique_ptr<qz_batch> b(new qz_batch());
        for (int i = 0; i < n; ++i) {
            if (!hs[i]) throw std::runtime_error("qz_batch_create: null run");
            if (hs[i]->s->m1 != hs[0]->s->m1) throw std::runtime_error("qz_batch_create: runs differ in modes");
            for (int j = 0; j < i; ++j)
                if (hs[j] == hs[i]) throw std::runtime_error("qz_batch_create: a run appears twice");
            b->sims.push_back(hs[i]);
        }
        b->use_graph = use_graph != 0;
        qz::hip_check(hipStreamCreate(&b->st), "stream");
        qz::hip_check(hipStreamCreate(&b->aux), "stream");
        for (auto& x : b->br) qz::hip_check(hipStreamCreate(&x), "stream");
        for (int f = 0; f < 2; ++f) {
            qz::hip_check(hipEventCreateWithFlags(&b->fork[f], hipEventDisableTiming), "event");
            for (auto& e : b->join[f]) qz::hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
        }
        qz::hip_check(hipEventCreateWithFlags(&b->mfork, hipEventDisableTiming), "event");
        qz::hip_check(hipEventCreateWithFlags(&b->mjoin, hipEventDisableTiming), "event");
        qz::hip_check(hipEventCreate(&b->e0), "event");
        qz::hip_check(hipEventCreate(&b->e1), "event");
        *out = b.release();
    })
}

extern "C" void qz_batch_destroy(qz_batch* b) { delete b; }

extern "C" int qz_batch_set_workers(qz_batch* b, int max_workers) {
    if (!b || max_workers < 0) return qz_fail(QZ_EINVAL, "qz_batch_set_workers: null batch or negative count");
    QZ_SIM_TRY({
        b->drop_graph();
        b->max_workers = max_workers;
    })
}

// capture of one step of every run: the per-run work (preparation, then beam push) on branch streams forked
// from and joined back to the batch stream around the sweep launches, so that runs overlap in the graph
// (each run's stream is switched to its branch meanwhile)
static void capture_batch(qz_batch* b) {
    const int n = static_cast<int>(b->sims.size());
    std::vector<qz::Simulation*> runs(n);
    for (int i = 0; i < n; ++i) runs[i] = b->sims[i]->s.get();
    b->plan = qz::plan_mixed(runs.data(), n, b->max_workers);   // the arguments are the same every step
    qz::upload_mixed(b->plan, b->dargs, b->st, true);
    std::vector<hipStream_t> saved(n);
    for (int i = 0; i < n; ++i) saved[i] = b->sims[i]->s->stream;
    const int nb = std::min(n, qz_batch::kBranches);
    auto per_run = [&](int phase, auto&& work) {   // fork, per-run work on the branches, join
        qz::hip_check(hipEventRecord(b->fork[phase], b->st), "fork");
        for (int j = 0; j < nb; ++j) qz::hip_check(hipStreamWaitEvent(b->br[j], b->fork[phase], 0), "fork");
        for (int i = 0; i < n; ++i) {
            b->sims[i]->s->stream = b->br[i % nb];
            work(*b->sims[i]->s);
            b->sims[i]->s->stream = saved[i];
        }
        for (int j = 0; j < nb; ++j) {
            qz::hip_check(hipEventRecord(b->join[phase][j], b->br[j]), "join");
            qz::hip_check(hipStreamWaitEvent(b->st, b->join[phase][j], 0), "join");
        }
    };
    qz::hip_check(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal), "begin capture");
    try {
        per_run(0, [](qz::Simulation& s) { s.prepare_sweep(false); });
        qz::enqueue_mixed(b->plan, b->dargs.p, b->st, b->aux, b->mfork, b->mjoin);
        per_run(1, [](qz::Simulation& s) { s.enqueue_push(false); });
    } catch (...) {
        for (int i = 0; i < n; ++i) b->sims[i]->s->stream = saved[i];
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(b->st, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    qz::hip_check(hipStreamEndCapture(b->st, &b->graph), "end capture");
    qz::hip_check(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "graph instantiate");
}

extern "C" int qz_batch_step(qz_batch* b, double* step_ms) {
    if (!b) return qz_fail(QZ_EINVAL, "qz_batch_step: null batch");
    QZ_SIM_TRY({
        const int n = static_cast<int>(b->sims.size());
        bool replay = b->use_graph;
        for (int i = 0; i < n; ++i) {
            const qz::Simulation& s = *b->sims[i]->s;
            if (s.done) throw std::runtime_error("qz_batch_step: a run has ended (time.steps reached)");
            replay = replay && s.capturable();
        }
        if (!replay) {   // start-up, last or output steps: the eager batched step
            const int rc = step_batched(b->sims.data(), n, step_ms, b->max_workers);
            if (rc != QZ_OK) throw std::runtime_error(qz_last_error());
            return QZ_OK;
        }
        if (!b->exec) capture_batch(b);
        qz::hip_check(hipEventRecord(b->e0, b->st), "event");
        qz::hip_check(hipGraphLaunch(b->exec, b->st), "graph launch");
        qz::hip_check(hipEventRecord(b->e1, b->st), "event");
        qz::hip_check(hipEventSynchronize(b->e1), "sync");
        float ms = 0;   // every multi-workgroup run's error flag is checked by its finish_sweep
        qz::hip_check(hipEventElapsedTime(&ms, b->e0, b->e1), "elapsed");
        ++b->replays;
        if (step_ms) *step_ms = ms;
        for (int i = 0; i < n; ++i) {   // the host side of the step (Simulation::finish_sweep, after_sweep)
            qz::Simulation& s = *b->sims[i]->s;
            s.finish_sweep(ms, false);
            std::vector<unsigned long long> cnt;
            s.counters.download(cnt, s.stream);
            s.beam_pushed = static_cast<double>(cnt[21]);
            s.ms_beam = 0;
            s.t += s.dt;
            s.dt_prev = s.dt;
            ++s.step_n;
        }
    })
}
extern "C" long qz_batch_replays(const qz_batch* b) { return b ? b->replays : -1; }

namespace {
// The run's sweep takes G workgroups from now on (a multi-workgroup workspace for G > 1).
void set_groups(qz::Simulation& s, int G) {
    if (G == s.groups && (G == 1 || s.mws.p)) return;
    s.groups = G;
    s.mws.release();
    if (G > 1) s.mws.alloc(qz_sweep_multi_workspace_doubles(s.M, qz::mobile_particles(s), G, s.m1 ? 1 : 0));
}

// Whether the run's sweep can take G workgroups.
bool groups_usable(const qz::Simulation& s, int G, int ncu) {
    const int np = qz::mobile_particles(s);
    if (G < 1 || G > ncu || np > G * 4096) return false;
    return G > 1 || qz_sweep_fits_one_group(s.m1, s.grid.N, np);
}

// Times the first nslices slices of the coming step's sweep with each candidate workgroup count, the best of
// `reps` launches each (ms_out[i] = -1: a count the run cannot use); the run keeps its workgroup count.
// Before the first step only: the samples overwrite the plasma arrays and the first rows of fld and diag,
// which the step reloads and rewrites.
void probe_groups(qz::Simulation& s, const int* cand, int ncand, int nslices, int reps, double* ms_out) {
    if (s.swept) throw std::runtime_error("sweep probe: probe before the first step");
    struct Events {
        hipEvent_t e[2] = {nullptr, nullptr};
        ~Events() { for (auto x : e) if (x) (void)hipEventDestroy(x); }
    } ev;
    qz::hip_check(hipEventCreate(&ev.e[0]), "event");
    qz::hip_check(hipEventCreate(&ev.e[1]), "event");
    const int keep = s.groups, ncu = qz::device_cus();
    for (int i = 0; i < ncand; ++i) {
        ms_out[i] = -1;
        if (!groups_usable(s, cand[i], ncu)) continue;
        set_groups(s, cand[i]);
        s.prepare_sweep(false);
        qz_sweep_args a = s.sweep_args;
        a.k_end = std::min(nslices, s.K);
        a.stats = nullptr;
        for (int r = 0; r < reps; ++r) {
            qz::hip_check(hipEventRecord(ev.e[0], s.stream), "event");
            qz::launch_sweeps(s.d_args, &a, 1, s.stream);
            qz::hip_check(hipEventRecord(ev.e[1], s.stream), "event");
            qz::hip_check(hipEventSynchronize(ev.e[1]), "sync");
            qz::check_multi_sweep(s.mws, s.groups, s.stream);
            float ms = 0;
            qz::hip_check(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]), "elapsed");
            if (ms_out[i] < 0 || ms < ms_out[i]) ms_out[i] = ms;
        }
    }
    set_groups(s, keep);
    qz::hip_check(hipStreamSynchronize(s.stream), "sync");
}

// Makespan of list scheduling: jobs (longest first) each to the earliest free of w workers.
double list_makespan(const std::vector<double>& jobs_desc, int w) {
    if (jobs_desc.empty()) return 0;
    if (w < 1) return 1e300;
    std::priority_queue<double, std::vector<double>, std::greater<double>> free_at;
    for (int i = 0; i < w; ++i) free_at.push(0.0);
    double end = 0;
    for (double t : jobs_desc) {
        const double s = free_at.top();
        free_at.pop();
        free_at.push(s + t);
        end = std::max(end, s + t);
    }
    return end;
}
}  // namespace

namespace {
// The batch autotuner's plan (qz_batch_balance, qz_batch_plan): est[i * ncand + c] is run i's estimated sweep
// with cand[c] workgroups (< 0: unusable).  A plan for a target T gives every run whose one-workgroup estimate
// exceeds T the fewest workgroups that bring it under T and list-schedules the other runs, longest first, on the
// compute units left; its cost is the larger of T and that makespan.  Every estimate is tried as T; the cheapest
// plan goes to G, its cost to cost[0] and the cost of one workgroup per run to cost[1] (-1: some run cannot).
void plan_groups(int n, const int* cand, int ncand, const double* est, int ncu, int* G, double* cost) {
    int c1 = -1;
    for (int c = 0; c < ncand; ++c)
        if (cand[c] == 1) c1 = c;
    auto E = [&](int i, int c) { return est[static_cast<size_t>(i) * ncand + c]; };
    auto one = [&](int i) { return c1 >= 0 ? E(i, c1) : -1.0; };
    auto plan_cost = [&](double T, std::vector<int>& g) -> double {
        int used = 0;
        double top = 0;
        std::vector<double> singles;
        for (int i = 0; i < n; ++i) {
            if (one(i) >= 0 && one(i) <= T) {
                g[i] = 1;
                singles.push_back(one(i));
                continue;
            }
            int pick = -1;
            for (int c = 0; c < ncand; ++c)
                if (cand[c] > 1 && E(i, c) >= 0 && E(i, c) <= T && (pick < 0 || cand[c] < cand[pick])) pick = c;
            if (pick < 0) return 1e300;
            g[i] = cand[pick];
            used += cand[pick];
            top = std::max(top, E(i, pick));
        }
        const int free_cus = ncu - used;   // the one launch of qz::plan_mixed
        if (used > ncu || (!singles.empty() && free_cus < 1)) return 1e300;
        std::sort(singles.begin(), singles.end(), std::greater<double>());
        const int w = std::min(static_cast<int>(singles.size()), free_cus);
        return std::max(top, list_makespan(singles, w));
    };
    std::vector<double> targets;
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < ncand; ++c)
            if (E(i, c) >= 0) targets.push_back(E(i, c));
    std::sort(targets.begin(), targets.end());
    targets.erase(std::unique(targets.begin(), targets.end()), targets.end());
    double best = 1e300;
    std::vector<int> g(n), bestg(n);
    for (double T : targets) {
        const double c = plan_cost(T, g);
        if (c < best * (1 - 1e-9)) {
            best = c;
            bestg = g;
        }
    }
    if (best >= 1e299) throw std::runtime_error("batch plan: no plan fits the compute units");
    double single_cost = -1;
    {
        std::vector<double> all;
        bool ok = true;
        for (int i = 0; i < n; ++i) {
            ok = ok && one(i) >= 0;
            all.push_back(one(i));
        }
        if (ok) {
            std::sort(all.begin(), all.end(), std::greater<double>());
            single_cost = list_makespan(all, std::min(n, ncu));
        }
    }
    for (int i = 0; i < n; ++i) G[i] = bestg[i];
    cost[0] = best;
    cost[1] = single_cost;
}
}  // namespace

// Autotuner of the sweep's workgroups per run (quarz_amd.h): every candidate count sweeps the first nslices
// slices of the coming step; the fastest is kept.
extern "C" int qz_sim_tune_groups(qz_sim* h, const int* cand, int ncand, int nslices, double* ms_out) {
    if (!h || !cand || ncand < 1 || nslices < 1 || !ms_out) return qz_fail(QZ_EINVAL, "qz_sim_tune_groups: bad arguments");
    QZ_SIM_TRY({
        qz::Simulation& s = *h->s;
        probe_groups(s, cand, ncand, nslices, 1, ms_out);
        int best = -1;
        for (int i = 0; i < ncand; ++i)
            if (ms_out[i] >= 0 && (best < 0 || ms_out[i] < ms_out[best])) best = i;
        if (best < 0) throw std::runtime_error("qz_sim_tune_groups: no candidate fits this run");
        set_groups(s, cand[best]);
    })
}

// Autotuner of a batch's workgroups per run (quarz_amd.h).  Runs of one shape (modes, cells, plasma particles)
// share a time per slice; the shape's run with the most slices is probed over its first nslices slices with
// every candidate count that keeps the smallest particle share (best of two launches).  A run's sweep is then
// estimated at (time per slice) x (its slices), and plan_groups picks the counts, which the runs then take.
extern "C" int qz_batch_balance(qz_batch* b, const int* cand, int ncand, int nslices, int* groups_out,
                                double* est_ms) {
    if (!b || !cand || ncand < 1 || nslices < 1 || !groups_out || !est_ms)
        return qz_fail(QZ_EINVAL, "qz_batch_balance: bad arguments");
    QZ_SIM_TRY({
        const int n = static_cast<int>(b->sims.size()), ncu = qz::device_cus();
        std::map<std::array<int, 3>, std::vector<int>> shapes;   // (modes, cells, particles) -> runs
        for (int i = 0; i < n; ++i) {
            const qz::Simulation& s = *b->sims[i]->s;
            if (s.swept) throw std::runtime_error("qz_batch_balance: balance before the first step");
            shapes[{s.m1 ? 1 : 0, s.grid.N, qz::mobile_particles(s)}].push_back(i);
        }
        // est[i * ncand + c]: run i's sweep with cand[c] workgroups (ms), < 0 where unusable
        std::vector<double> est(static_cast<size_t>(n) * ncand, -1.0), per(ncand);
        int c1 = -1;
        for (int c = 0; c < ncand; ++c)
            if (cand[c] == 1) c1 = c;
        for (const auto& kv : shapes) {
            int rep = kv.second[0];
            for (int i : kv.second)
                if (b->sims[i]->s->K > b->sims[rep]->s->K) rep = i;
            qz::Simulation& r = *b->sims[rep]->s;
            // the multi-workgroup runs share one launch, whose particle share per workgroup follows the largest
            // particles-per-thread of its runs: only counts at the smallest (1024 particles or fewer per
            // workgroup) are offered, so that every run keeps its own share
            std::vector<int> ok_cand;
            std::vector<int> at;
            for (int c = 0; c < ncand; ++c)
                if (cand[c] == 1 || qz_sweep_multi_pt(kv.first[2], cand[c]) == 2) {
                    ok_cand.push_back(cand[c]);
                    at.push_back(c);
                }
            std::vector<double> got(ok_cand.size(), -1);
            if (!ok_cand.empty())
                probe_groups(r, ok_cand.data(), static_cast<int>(ok_cand.size()), nslices, 2, got.data());
            std::fill(per.begin(), per.end(), -1.0);
            for (size_t j = 0; j < at.size(); ++j) per[at[j]] = got[j];
            const double sampled = std::min(nslices, r.K);
            for (int i : kv.second) {
                qz::Simulation& s = *b->sims[i]->s;
                for (int c = 0; c < ncand; ++c)
                    est[static_cast<size_t>(i) * ncand + c] = per[c] < 0 ? -1 : per[c] / sampled * s.K;
                s.slice_ms = c1 >= 0 && per[c1] > 0 ? per[c1] / sampled : 0;
            }
        }
        std::vector<int> G(n);
        double cost[2];
        plan_groups(n, cand, ncand, est.data(), ncu, G.data(), cost);
        b->drop_graph();
        for (int i = 0; i < n; ++i) {
            set_groups(*b->sims[i]->s, G[i]);
            groups_out[i] = G[i];
        }
        est_ms[0] = cost[0];
        est_ms[1] = cost[1];
    })
}

extern "C" int qz_batch_plan(int n, const int* cand, int ncand, const double* est, int ncu, int* groups_out,
                             double* est_ms) {
    if (n < 1 || !cand || ncand < 1 || !est || ncu < 1 || !groups_out || !est_ms)
        return qz_fail(QZ_EINVAL, "qz_batch_plan: bad arguments");
    QZ_SIM_TRY(plan_groups(n, cand, ncand, est, ncu, groups_out, est_ms))
}

// Autotuner step for one beam kernel (quarz_amd.h): candidates timed on scratch copies of the largest beam
extern "C" int qz_sim_autotune(qz_sim* h, int kernel, const int* cand, int ncand, int reps, double* ms_out) {
    if (!h || !cand || ncand < 1 || reps < 1 || !ms_out) return qz_fail(QZ_EINVAL, "qz_sim_autotune: bad arguments");
    if (kernel < 0 || kernel >= QZ_K_COUNT) return qz_fail(QZ_EINVAL, "qz_sim_autotune: unknown kernel");
    QZ_SIM_TRY({
        qz::Simulation& s = *h->s;
        const bool push = kernel == QZ_K_BEAM_ADVANCE;
        const qz::Beam* bm = nullptr;
        for (const auto& b : s.beams)
            if ((!push || !b->rigid) && (!bm || b->Np > bm->Np)) bm = b.get();
        for (int i = 0; i < ncand; ++i) ms_out[i] = 0;
        if (!bm || bm->Np == 0) return QZ_OK;   // nothing to time: the current shape stays
        const size_t np = static_cast<size_t>(bm->Np);
        qz::DevBuf<qz::Real> a[7];
        for (int c = 0; c < 7; ++c) {
            a[c].alloc(np);
            qz::hip_check(hipMemcpyAsync(a[c].p, bm->a[c].p, 8 * np, hipMemcpyDeviceToDevice, s.stream), "scratch copy");
        }
        const qz_beam_soa soa{a[0].p, a[1].p, a[2].p, a[3].p, a[4].p, a[5].p, a[6].p};
        qz::DevBuf<qz::Real> src;
        if (!push) src.alloc(s.bsrc.n);
        const qz_radial_grid g = s.grid.view();
        const int m1 = s.m1 ? 1 : 0;
        auto run = [&]() {
            if (push) {
                qz::qz_check(qz_beam_advance(&g, &s.box, m1, bm->q / bm->m, bm->pusher, 0, bm->Np, &soa, s.fld.p, s.NCF,
                                             s.dt, s.dt, 1, s.stream));
            } else {
                qz::qz_check(qz_zero_async(src.p, sizeof(qz::Real) * src.n, s.stream));
                qz::qz_check(qz_beam_deposit(&g, &s.box, m1, bm->q, bm->Np, &soa, src.p, s.NCB, s.dep_work.p, s.stream));
            }
        };
        struct Events {
            hipEvent_t e[2] = {nullptr, nullptr};
            ~Events() { for (auto x : e) if (x) (void)hipEventDestroy(x); }
        } ev;
        qz::hip_check(hipEventCreate(&ev.e[0]), "event");
        qz::hip_check(hipEventCreate(&ev.e[1]), "event");
        int best = -1;
        for (int i = 0; i < ncand; ++i) {
            if (qz_set_block(kernel, cand[i]) != QZ_OK) throw std::runtime_error(qz_last_error());
            run();   // warm-up
            qz::hip_check(hipEventRecord(ev.e[0], s.stream), "event");
            for (int r = 0; r < reps; ++r) run();
            qz::hip_check(hipEventRecord(ev.e[1], s.stream), "event");
            qz::hip_check(hipEventSynchronize(ev.e[1]), "sync");
            float ms = 0;
            qz::hip_check(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]), "elapsed");
            ms_out[i] = ms / reps;
            if (best < 0 || ms_out[i] < ms_out[best]) best = i;
        }
        if (qz_set_block(kernel, cand[best]) != QZ_OK) throw std::runtime_error(qz_last_error());
        qz::hip_check(hipStreamSynchronize(s.stream), "sync");
    })
}

extern "C" int qz_sim_run_pipelined(qz_sim* h, int nrun, int nchunks, int depth, double* out2) {
    QZ_NEED(h, "qz_sim_run_pipelined");
    QZ_SIM_TRY({
        double psl = 0, pushed = 0;
        h->s->run_pipelined(nrun, nchunks, depth, psl, pushed);
        if (out2) { out2[0] = psl; out2[1] = pushed; }
    })
}
extern "C" int qz_sim_step_fields(qz_sim* h) {
    QZ_NEED(h, "qz_sim_step_fields");
    QZ_SIM_TRY(h->s->step_fields())
}
extern "C" int qz_sim_prepare_sweep(qz_sim* h, qz_sweep_args* out) {
    QZ_NEED(h && out, "qz_sim_prepare_sweep");
    QZ_SIM_TRY({
        auto& s = *h->s;
        s.prepare_sweep();
        qz::hip_check(hipStreamSynchronize(s.stream), "sync");
        *out = s.sweep_args;
    })
}
extern "C" int qz_sim_push_beams(qz_sim* h) {
    QZ_NEED(h, "qz_sim_push_beams");
    QZ_SIM_TRY({
        auto& s = *h->s;
        const bool last = (s.step_n >= s.nsteps);
        s.push_beams();
        if (last) s.done = true;
        ++s.step_n;
    })
}
extern "C" long qz_sim_get(qz_sim* h, const char* what, double* out, long cap) {
    if (!h || !what) {
        qz_fail(QZ_EINVAL, "qz_sim_get: null argument");
        return -1;
    }
    try {
        auto& s = *h->s;
        const std::string w(what);
        std::vector<double> v;
        if (w == "fld") s.fld.download(v, s.stream);
        else if (w == "chi") {
            if (!s.chi_rec.p) return -1;
            s.chi_rec.download(v, s.stream);
        } else if (w == "diag") s.diag.download(v, s.stream);
        else if (w == "bsrc") s.bsrc.download(v, s.stream);
        else if (w == "nodes") v = s.grid.r;
        else if (w.rfind("beam", 0) == 0) {
            const int b = std::atoi(w.c_str() + 4);
            const std::string c = w.substr(w.find('.') + 1);
            if (b < 0 || b >= static_cast<int>(s.beams.size())) return -1;
            static const char* names[7] = {"x", "y", "px", "py", "pz", "xi", "w"};
            int ci = -1;
            for (int i = 0; i < 7; ++i) if (c == names[i]) ci = i;
            if (ci < 0) return -1;
            s.beams[b]->a[ci].download(v, s.stream);
            v.resize(static_cast<size_t>(s.beams[b]->Np));
        } else if (w.rfind("sp", 0) == 0) {
            const int b = std::atoi(w.c_str() + 2);
            const std::string c = w.substr(w.find('.') + 1);
            if (b < 0 || b >= static_cast<int>(s.species.size())) return -1;
            auto& S = *s.species[b];
            const qz::DevBuf<double>* src = c == "x" ? &S.x : c == "y" ? &S.y : c == "px" ? &S.px : c == "py" ? &S.py
                                          : c == "dl" ? &S.dl : c == "w" ? &S.w : nullptr;
            if (!src) return -1;
            src->download(v, s.stream);
            v.resize(static_cast<size_t>(S.Np));
        } else {
            return -1;
        }
        const long n = static_cast<long>(v.size());
        if (out && cap >= n) std::copy(v.begin(), v.end(), out);
        return n;
    } catch (const std::exception& e) {
        qz_fail(QZ_EINVAL, e.what());
        return -1;
    }
}
extern "C" int qz_sim_set_beam_profile(qz_sim* h, int b, const double* values, int nseg) {
    if (!h || !values || nseg < 1) return qz_fail(QZ_EINVAL, "qz_sim_set_beam_profile: bad argument");
    QZ_SIM_TRY({
        auto& s = *h->s;
        if (b < 0 || b >= static_cast<int>(s.beams.size())) throw std::runtime_error("no such beam");
        s.beams[b]->set_profile(std::vector<double>(values, values + nseg), s.stream);
    })
}
extern "C" int qz_sim_spread_grad(qz_sim* h, int b, int every, double* J, double* grad, int nseg, double* ms2) {
    if (!h || !J || !grad) return qz_fail(QZ_EINVAL, "qz_sim_spread_grad: null argument");
    QZ_SIM_TRY({
        auto& s = *h->s;
        if (b < 0 || b >= static_cast<int>(s.beams.size())) throw std::runtime_error("no such beam");
        if (nseg != static_cast<int>(s.beams[b]->pwl.size()))
            throw std::runtime_error("qz_sim_spread_grad: the profile has " + std::to_string(s.beams[b]->pwl.size()) +
                                     " values");
        std::vector<double> g;
        s.spread_value_grad(b, every, *J, g);
        std::copy(g.begin(), g.end(), grad);
        if (ms2) { ms2[0] = s.ms_adj_forward; ms2[1] = s.ms_adj_reverse; }
    })
}
extern "C" int qz_sim_adjoint_ticks(qz_sim* h, double* out) {
    if (!h || !out) return qz_fail(QZ_EINVAL, "qz_sim_adjoint_ticks: null argument");
    for (int i = 0; i < QZ_ADJ_NPHASE; ++i) out[i] = h->s->adj ? h->s->adj->last_ticks[i] : 0.0;
    return QZ_OK;
}
// chi_0 of every slice of the coming sweeps into a (K, M) device array (the plasma susceptibility the laser
// envelope's advance_slice takes, src/Simulation.cpp:735); on = 0 stops and frees it
extern "C" int qz_sim_record_chi(qz_sim* h, int on) {
    QZ_NEED(h, "qz_sim_record_chi");
    QZ_SIM_TRY({
        auto& s = *h->s;
        if (on) {
            if (!s.chi_rec.p) {
                s.chi_rec.alloc(static_cast<size_t>(s.K) * s.M);
                qz::qz_check(qz_zero_async(s.chi_rec.p, sizeof(double) * s.chi_rec.n, s.stream));
                qz::hip_check(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
            }
        } else {
            qz::hip_check(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
            s.chi_rec.release();
        }
    })
}
// the device address of that array (null when not recording), valid until recording stops
extern "C" int qz_sim_chi_device(qz_sim* h, double** ptr) {
    QZ_NEED(h && ptr, "qz_sim_chi_device");
    *ptr = h->s->chi_rec.p;
    return QZ_OK;
}

extern "C" int qz_sim_profile_size(qz_sim* h, int b) {
    if (!h) return qz_fail(QZ_EINVAL, "qz_sim_profile_size: null run");
    if (b < 0 || b >= static_cast<int>(h->s->beams.size())) return qz_fail(QZ_EINVAL, "qz_sim_profile_size: no such beam");
    return static_cast<int>(h->s->beams[b]->pwl.size());
}
extern "C" int qz_sim_snapshot(qz_sim* h) {
    if (!h) return qz_fail(QZ_EINVAL, "qz_sim_snapshot: null run");
    QZ_SIM_TRY(h->s->snapshot())
}
extern "C" int qz_sim_rewind(qz_sim* h) {
    if (!h) return qz_fail(QZ_EINVAL, "qz_sim_rewind: null run");
    QZ_SIM_TRY(h->s->rewind())
}
extern "C" int qz_sim_beam_sums(qz_sim* h, int b, double* out17, double* refs) {
    QZ_NEED(h && out17, "qz_sim_beam_sums");
    QZ_SIM_TRY({
        auto& s = *h->s;
        if (b < 0 || b >= static_cast<int>(s.beams.size())) throw std::runtime_error("no such beam");
        const auto S = s.beams[b]->sums(s.stream);
        std::copy(S.begin(), S.end(), out17);
        if (refs) { refs[0] = s.beams[b]->gref; refs[1] = s.beams[b]->xiref; }
    })
}
extern "C" int qz_sim_beam_diag(qz_sim* h, int b, double* out15) {
    QZ_NEED(h && out15, "qz_sim_beam_diag");
    QZ_SIM_TRY({
        auto& s = *h->s;
        if (b < 0 || b >= static_cast<int>(s.beams.size())) throw std::runtime_error("no such beam");
        const qz::BeamDiag d = s.beams[b]->from_sums(s.beams[b]->sums_device(s.stream), s.t);
        const double v[15] = {d.t, d.npart, d.charge, d.gamma_mean, d.gamma_rms, d.xi_mean, d.xi_rms, d.r_rms,
                              d.emit_nx, d.x_mean, d.y_mean, d.emit_ny, d.sigma_x, d.sigma_y, double(d.alive)};
        std::copy(v, v + 15, out15);
    })
}
extern "C" int qz_sim_timing(const qz_sim* h, double* out2) {
    if (!h || !out2) return qz_fail(QZ_EINVAL, "qz_sim_timing: null argument");
    out2[0] = h->s->ms_sweep;
    out2[1] = h->s->ms_beam;
    return QZ_OK;
}
extern "C" int qz_sim_stats(const qz_sim* h, double* out7) {
    if (!h || !out7) return qz_fail(QZ_EINVAL, "qz_sim_stats: null argument");
    const auto& s = *h->s;
    out7[0] = s.particle_slices;
    out7[1] = s.beam_pushed;
    out7[2] = s.beam_deposited;
    out7[3] = s.ms_sweep;
    out7[4] = s.ms_beam_dep;
    out7[5] = s.ms_beam;
    out7[6] = s.sweep_launches;
    return QZ_OK;
}
extern "C" long qz_sim_pipeline_timeline(const qz_sim* h, double* out, long cap, int* nchunks) {
    if (!h) return qz_fail(QZ_EINVAL, "qz_sim_pipeline_timeline: null run");
    const std::vector<double>& t = h->s->pipe_timeline;
    const long n = static_cast<long>(t.size() / 4);
    if (out)
        for (long i = 0; i < std::min(n, cap); ++i)
            for (int m = 0; m < 4; ++m) out[4 * i + m] = t[4 * i + m];
    if (nchunks) *nchunks = h->s->pipe_chunks;
    return n;
}

extern "C" int qz_sim_phase_ticks(const qz_sim* h, double* out) {
    if (!h || !out) return qz_fail(QZ_EINVAL, "qz_sim_phase_ticks: null argument");
    for (int i = 0; i < QZ_NPHASE; ++i) out[i] = h->s->phase_ticks[i];
    return QZ_OK;
}
extern "C" int qz_sim_groups(const qz_sim* h) {
    if (!h) return qz_fail(QZ_EINVAL, "qz_sim_groups: null argument");
    return h->s->groups;
}
extern "C" long qz_sim_lost(const qz_sim* h) {
    if (!h) {
        qz_fail(QZ_EINVAL, "qz_sim_lost: null argument");
        return -1;
    }
    return h->s->lost_last;
}
extern "C" double qz_sim_time(const qz_sim* h) {
    if (!h) {
        qz_fail(QZ_EINVAL, "qz_sim_time: null argument");
        return std::nan("");
    }
    return h->s->t;
}
extern "C" int qz_sim_write_fields(qz_sim* h, const char* path) {
    QZ_NEED(h && path, "qz_sim_write_fields");
    QZ_SIM_TRY(h->s->write_fields(path))
}
