// Host side of the quasi-static step: the classes of the reference's interface (RadialGrid, FieldSolver,
// PlasmaSpecies, Beam, Simulation; src/*.hpp), with device arrays in HBM and the hot path in the gfx950
// kernels behind the C-ABI of include/quarz_amd.h.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "../../../include/quarz_amd.h"
#include "config.hpp"
#include "../qz_internal.hpp"

namespace qz {

using Real = double;

// ---------------------------------------------------------------------------- device buffer
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count);
    void release();
    void upload(const std::vector<T>& h, hipStream_t st);
    void download(std::vector<T>& h, hipStream_t st) const;
};

// ---------------------------------------------------------------------------- RadialGrid (src/RadialGrid.*)
struct RadialGrid {
    int N = 0;
    Real R = 0;
    std::vector<Real> r, h, rmid, V, inv_h, inv_V;
    DevBuf<Real> d_r, d_inv_h, d_inv_V;

    static std::vector<Real> make_uniform(Real R, Real dr);
    static std::vector<Real> make_regions(Real R, const std::vector<std::pair<double, double>>& regions, Real max_ratio);
    static std::vector<Real> read_file(const std::string& filename);
    void finalize(const std::vector<Real>& nodes, hipStream_t st);
    qz_radial_grid view() const { return {d_r.p, d_inv_h.p, d_inv_V.p, N, R}; }
};

// ---------------------------------------------------------------------------- FieldSolver (src/FieldSolver.*)
struct FieldSolver {
    enum { L0 = 0, L1D = 1, L1F = 2, L2D = 3 };
    std::vector<Real> a[4], b[4], c[4];
    Real flux_coef = 0;
    std::vector<Real> gm, g0, gp, ir;
    std::vector<Real> l0_m, l0_cp, l1_m, l1_cp;   // Thomas factors of the chi-free L0 and L1D
    DevBuf<Real> d_a[4], d_b[4], d_c[4], d_gm, d_g0, d_gp, d_ir, d_l0_m, d_l0_cp, d_l1_m, d_l1_cp;
    void init(const RadialGrid& g, hipStream_t st);
};

// ---------------------------------------------------------------------------- PlasmaSpecies (src/PlasmaSpecies.*)
struct PlasmaSpecies {
    std::string name;
    bool m1 = false;
    Real q = -1, m = 1;
    bool mobile = true;
    int ppc = 4, ntheta = 1;
    Real uth = 0, density = 1, rmax = 1e300;
    Real delta_min = 1e-3, max_qsa = 35;
    int ab_order = 3;
    Real ab_c[5] = {0, 0, 0, 0, 0};
    int Np = 0, cap = 0;
    std::vector<Real> hx, hy, hpx, hpy, hdl, hw;   // host copy of the last load
    DevBuf<Real> x, y, px, py, dl, w, hist;
    // The load does not depend on the step (uniform profile, the same seed every step): the first load is
    // kept in HBM and later loads are device copies of it.
    DevBuf<Real> t_x, t_y, t_px, t_py, t_dl, t_w;
    bool have_template = false;
    // A cold plasma (uth = 0) is loaded by the sweep kernel itself at the head of the box (qz_plasma_rings):
    // the rings loaded and the cos / sin table of their angles.
    int nring = 0;
    DevBuf<Real> d_cs;
    bool device_rings() const { return have_template && uth == 0; }

    void init(const Config& cfg, const std::string& nm, const RadialGrid& g, bool mode1);
    void load(const RadialGrid& g, Real z, unsigned long long seed, hipStream_t st);
    // rho - J_z and rho of the loaded particles, mode arrays (3 M each), host (src/PlasmaSpecies.cpp:480-506)
    void deposit_rho(const RadialGrid& g, std::vector<Real>& rhot, std::vector<Real>& rho) const;
    qz_species_state state() const {
        return {x.p, y.p, px.p, py.p, dl.p, w.p, hist.p, Np, cap, q, m};
    }
};

// ---------------------------------------------------------------------------- Beam (src/Beam.*)
double pwl_factor(const std::vector<double>& v, double u0, double u1, double u);

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
    int pusher = 0;
    bool started = false;
    double gref = 0, xiref = 0;
    int Np = 0;
    DevBuf<Real> a[7];   // x y px py pz xi w
    DevBuf<int> pstep;   // step that last pushed each particle (pipelined runs)
    // piecewise-linear weight profile of a flat-top beam (<name>.pwl_weights): values, user-xi range, the
    // flat-top macro-particle weight, and the sampled xi and gamma of every particle
    std::vector<double> pwl, xi_init, g_init;
    double pwl_u0 = 0, pwl_u1 = 0, pwl_wmac = 0;
    void init(const Config& cfg, const std::string& nm, bool mode1, hipStream_t st);
    void set_profile(const std::vector<double>& values, hipStream_t st);   // new weights from the profile
    void set_particles(const std::vector<double>& p7, hipStream_t st);
    qz_beam_soa soa() const { return {a[0].p, a[1].p, a[2].p, a[3].p, a[4].p, a[5].p, a[6].p}; }
    // the 17 sums of Beam::sums in the reference's index order (host copy of the particles): the checker's form
    std::array<double, 17> sums(hipStream_t st) const;
    // the same sums reduced on the device (qz_beam_sums_device, fixed tree order): the beam log and the
    // differentiable objective; only 17 doubles cross to the host
    std::array<double, 17> sums_device(hipStream_t st) const;
    BeamDiag from_sums(const std::array<double, 17>& S, Real t) const;
    mutable DevBuf<Real> sums_work;   // 17 results, then the reduction's block partials
};

// ---------------------------------------------------------------------------- Simulation (src/Simulation.*)
class Simulation {
public:
    Simulation(const std::string& deck, const std::string& overrides);
    ~Simulation();

    void step_fields();      // the quasi-static sweep of the current step n
    // step_fields before the slice loop; fills sweep_args.  With timed = false it records no events: all its
    // work is then device work on `stream` (capturable) once the background and the plasma template exist.
    void prepare_sweep(bool timed = true);
    void finish_sweep(float sweep_ms, bool timed = true);   // step_fields after the slice loop
    void push_beams();       // the beam push of step n (not on the last step)
    void enqueue_push(bool timed);       // the device work of push_beams (live count, kick and drift launches)
    bool capturable() const;             // the next step's device work is the same as the last one's
    void after_sweep();      // output, beam push and step counter after the sweep
    void step();             // one iteration of Simulation::run
    // nrun steps with up to `depth` steps in flight, every step split into nchunks slice ranges (see the
    // definition); bit-identical to nrun calls of step().  Returns the particle-slices and beam particles
    // pushed of all nrun steps.
    void run_pipelined(int nrun, int nchunks, int depth, double& psl_total, double& pushed_total);
    void write_fields(const std::string& path);
    void write_beam_output();   // one beams.txt line per beam: the beam log of the current step
    void write_axis(const std::string& path);   // axis_*.txt: on-axis fields of every slice
    // the beams and the step state as they are now, and back to them (repeated evaluations of one step)
    void snapshot();
    void rewind();
    // One step (m = 0) and the gradient of its objective J = gamma_rms / gamma_mean of beam b after the step with
    // respect to b's weight-profile values, by the adjoint: the sweep runs in segments of `every` slices with the
    // plasma state saved at each segment's start; the reverse sweep recomputes each segment from its
    // checkpoint, last segment first.  The run ends the call one step on, as after step().
    void spread_value_grad(int b, int every, double& J, std::vector<double>& grad);
    struct AdjointBufs {
        DevBuf<Real> ckpt, traj[2], ring, work, lam_fld, lam_bsrc, lam_gamma, lam_w, u, grad;
        DevBuf<Real> before[7];   // the beam before its push
        DevBuf<unsigned long long> ticks;   // the reverse sweep's phase clocks
        double last_ticks[QZ_ADJ_NPHASE] = {};
        int u_beam = -1;          // the beam whose sampled xi `u` holds
        // the reverse sweep runs on its own stream while the run's stream recomputes the next segment
        hipStream_t st2 = nullptr;
        std::vector<hipEvent_t> ev;   // per segment: recomputed, reversed
        AdjointBufs() = default;
        AdjointBufs(const AdjointBufs&) = delete;
        AdjointBufs& operator=(const AdjointBufs&) = delete;
        ~AdjointBufs();
    };
    std::unique_ptr<AdjointBufs> adj;
    float ms_adj_forward = 0, ms_adj_reverse = 0;   // the checkpointed sweep and the reverse part of the last call

    bool m1 = false;
    int picard = 1;
    RadialGrid grid;
    FieldSolver solver;
    qz_beam_grid box{};
    Real dt = 0, t = 0;
    int nsteps = 0, step_n = 0;
    bool done = false, have_prev = false, swept = false;
    double dt_prev = 0;
    unsigned long long seed = 1;
    bool neutralize = true;
    std::vector<std::unique_ptr<PlasmaSpecies>> species;
    std::vector<std::unique_ptr<Beam>> beams;
    int M = 0, K = 0, NCF = 0, NCB = 0, NCD = 0;
    DevBuf<Real> fld, bsrc, diag, bg, ws;
    DevBuf<Real> chi_rec;   // (K, M): chi_0 of every slice of the sweeps while recording (qz_sim_record_chi)
    DevBuf<int> lost;
    DevBuf<unsigned long long> counters;   // [0] particle-slices, [1..16] sweep phase ticks, [20] beam deposited, [21] pushed
    DevBuf<char> dep_work;
    qz_sweep_args sweep_args{};
    DevBuf<qz_sweep_args> d_args;   // this run's device copy of the sweep arguments (batched: every run's)
    std::vector<Real> bg_host;   // host side of the background upload (kept alive for the async copy)
    bool bg_ready = false;       // the background is the same every step: computed and uploaded once
    long lost_last = 0;
    float ms_sweep = 0, ms_beam = 0, ms_beam_dep = 0;
    int sweep_launches = 1;      // sweep kernel launches of the last step (pipelined: its slice ranges)
    std::vector<double> pipe_timeline;   // last run_pipelined: 4 marks (ms) per range launch, step-major
    int pipe_chunks = 0;
    int groups = 1;              // workgroups of the sweep (1: the one-workgroup kernel)
    DevBuf<Real> mws;            // groups > 1: the multi-workgroup workspace
    double slice_ms = 0;         // a batch's estimate of the one-workgroup sweep's time per slice (0: none)
    hipStream_t aux = nullptr;   // a batched step's launch stream for the multi-workgroup runs (made on first use)
    hipEvent_t aux_ev[2] = {};   // fork to and join from `aux`
    double particle_slices = 0, beam_pushed = 0, beam_deposited = 0;
    double phase_ticks[QZ_NPHASE] = {};
    hipStream_t stream = nullptr;
    hipEvent_t ev[6] = {};
    std::string outdir;
    int out_every = 0;
    bool out_axis = true;    // output.axis: axis_*.txt with every field output
    std::ofstream beamlog;   // outdir/beams.txt
    struct Snap {
        std::vector<std::array<DevBuf<Real>, 7>> beams;
        std::vector<char> started;
        std::vector<std::array<double, 2>> refs;     // gref, xiref
        std::vector<std::vector<double>> pwl;
        Real t = 0;
        int step_n = 0;
        bool done = false, have_prev = false, swept = false;
        double dt_prev = 0;
    };
    std::unique_ptr<Snap> snap;
};

}  // namespace qz
