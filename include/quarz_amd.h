/*
 * quarz_amd -- C-ABI of the MI355X (gfx950) quasi-static step of QUARZ.
 *
 * The reference has no plugin/FFI layer: its hot path is a set of C++ member functions dispatched
 * through Kokkos.  Each entry point below replaces one of them (reference file:line cited); the C++ host
 * classes of this library (Simulation, PlasmaSpecies, Beam, FieldSolver, RadialGrid, Config) call these
 * exactly where the reference calls the member functions.
 *
 * Conventions
 *   - Plain pointers and sizes only.  Device pointers are hipMalloc'd memory of the current device;
 *     `stream` is a hipStream_t (NULL: the default stream).  The library never frees caller memory.
 *   - Arrays are FP64, layouts as the reference: (xi, r, component) arrays LayoutRight with the component
 *     fastest (src/FieldStore.hpp:9-21), particle data SoA.
 *   - Every function returns 0 on success and a negative code on failure (bad extent, unsupported size,
 *     HIP error); qz_last_error() returns the message.  Kernels never fail: removed particles are data
 *     (w = 0), as in the reference.
 *   - Results are bit-identical to a serial (single-thread) CPU build of the reference arithmetic:
 *     deposits are summed per node in particle-index order, no multiply-add is fused, and divisions and
 *     square roots are correctly rounded.
 */
#ifndef QUARZ_AMD_H
#define QUARZ_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes */
#define QZ_OK 0
#define QZ_EINVAL (-1)    /* bad argument or extent */
#define QZ_ERANGE (-2)    /* size beyond what the kernels support */
#define QZ_EHIP (-3)      /* HIP runtime error */

/* Device view of the radial grid (GridD, src/RadialGrid.hpp:26-55). */
typedef struct qz_radial_grid {
    const double* r;      /* N+1 nodes */
    const double* inv_h;  /* N: 1 / h_j */
    const double* inv_V;  /* N+1: 1 / V_j (hat-function volumes) */
    int N;                /* cells */
    double R;             /* outer radius */
} qz_radial_grid;

/* Longitudinal box of the beams (BeamGrid, src/Beam.hpp:25-30). */
typedef struct qz_beam_grid {
    double xi_min, dxi;
    int nxi, k0, nloc;
    int ngp;              /* 1: nearest-slice shape in xi, 0: linear */
} qz_beam_grid;

/* Beam particles, SoA x y px py pz xi w (Beam::x_ ... w_, src/Beam.hpp:168). */
typedef struct qz_beam_soa {
    double *x, *y, *px, *py, *pz, *xi, *w;
} qz_beam_soa;

/* Mobile plasma species state (PlasmaSpecies, src/PlasmaSpecies.hpp:132-177). */
typedef struct qz_species_state {
    double *x, *y, *px, *py, *dl, *w;   /* np live slots; dl = Delta = gamma - p_z */
    double* hist;                       /* AB history (ab_order, 5, cap), particle fastest */
    int np, cap;
    double q, m;
} qz_species_state;

#define QZ_MAX_SPECIES 2

/* Cold plasma rings loaded at the head of the box by the sweep itself (PlasmaSpecies::load,
 * src/PlasmaSpecies.cpp:202-295): ring r (r = ppc j + sub) is the sub-annulus sub of cell j at its volume
 * centroid, with weight density pi (rb^2 - ra^2) / ntheta per particle; particle l of ring r sits at angle
 * 2 pi (l + (r % 2) / 2) / ntheta.  The first nring rings are loaded (those beyond plasma.rmax carry no weight). */
typedef struct qz_plasma_rings {
    int ppc, ntheta, nring;
    double density;
    const double* cs;     /* device: cos then sin of the 2 ntheta angles, turn 0 first (4 ntheta doubles) */
} qz_plasma_rings;
#define QZ_NPHASE 16      /* sweep phase clocks: see qz_sim_phase_ticks */

/* Everything one quasi-static sweep reads and writes (Simulation::step_fields, src/Simulation.cpp:636-795). */
typedef struct qz_sweep_args {
    qz_radial_grid grid;
    int modes;                 /* 0: m = 0; 1: m = 0 and 1 */
    int picard;                /* Picard iterations of the chi_1 coupling (m = 1) */
    int nslices;               /* K slices, head first */
    int k_begin, k_end;        /* slices of this launch, [k_begin, k_end); k_end = 0 means K.  A launch that starts
                                  past 0 continues the state the launch ending at k_begin left: plasma arrays, AB
                                  history and the B+ warm start in ws (B+ starts at zero only at slice 0) */
    int ab_order;              /* Adams-Bashforth order 1..5 */
    int ngp;                   /* beams.xi_shape = ngp */
    double dxi, delta_min, max_qsa;
    double ab_c[5];            /* AB weights of every push of the sweep (uniform steps) */
    /* operators L0, L1D, L1F, L2D: tridiagonal rows a, b, c (FieldSolver, src/FieldSolver.cpp:43-109) */
    const double* op_a[4];
    const double* op_b[4];
    const double* op_c[4];
    double flux_coef;          /* g coefficient of the flux row of L1F */
    const double *grad_m, *grad_0, *grad_p, *inv_r;   /* gradient stencils, 1/r (FieldSolver.cpp:157-176) */
    /* Thomas factors of the chi-free L0 and L1D: m (M) and c' (M) as Thomas forms them (Tridiag.cpp:114-121) */
    const double *fac_l0_m, *fac_l0_cp, *fac_l1_m, *fac_l1_cp;
    const double* bsrc;        /* (K, M, ncb) beam sources */
    int ncb;
    const double* bg;          /* background rho - J_z (3 M: mode 0, 1r, 1i) then rho (3 M) */
    double* fld;               /* (K, M, ncf) fields for the beams */
    int ncf;
    double* diag;              /* (K, M, ncd) psi, n_e, n_i */
    int ncd;
    int nspecies;              /* mobile species (<= QZ_MAX_SPECIES) */
    qz_species_state sp[QZ_MAX_SPECIES];
    int* lost;                 /* one counter per species (particles removed as trapped) */
    double* ws;                /* per-run workspace of qz_sweep_workspace_doubles(M) doubles */
    unsigned long long* stats; /* optional: stats[0] += live particles of every slice (particle-slices),
                                  stats[1..QZ_NPHASE] += wall-clock ticks of the sweep phases */
    int groups;                /* workgroups of the run: 0 or 1 -- one workgroup (at most 4096 plasma particles and a
                                  grid whose solve sets fit one compute unit); G > 1 -- G cooperating workgroups of at
                                  most 4096 particles each, for larger runs (no phase clocks) */
    double* mws;               /* G > 1: the run's workspace of qz_sweep_multi_workspace_doubles(M, np, G, modes)
                                  doubles (per-slice lists, counts and the barrier of the workgroups) */
    int load;                  /* 1: a launch that starts at slice 0 first loads every mobile species' cold rings
                                  (rings[s]) into its arrays and clears its AB history; 0: the arrays hold the load */
    qz_plasma_rings rings[QZ_MAX_SPECIES];
    double* traj;              /* optional, one workgroup: (k_end - k_begin) x 6 x np doubles, np = the species' counts
                                  summed: x, y, px, py, Delta, w of every particle (species 0 then 1) at the start
                                  of each slice of the launch (the adjoint's recomputed segments) */
    double* chi_out;           /* optional: (K, M) doubles; row k receives slice k's plasma susceptibility chi_0 (the
                                  mode-0 chi of the B+ solves) as the sweep forms it, the chi that advance_slice of
                                  the laser envelope takes (src/Simulation.cpp:735) */
} qz_sweep_args;

const char* qz_last_error(void);

/* Library / device information: gcnArchName and the number of CUs of the current device. */
int qz_device_info(char* arch, int arch_len, int* n_cu);

/* Batched Thomas solve of `batch` systems of size M (TridiagSolver::solve, Thomas branch,
 * src/Tridiag.cpp:111-126): element i of system s at s * sys_stride + i * elem_stride of a, b, c, d and x
 * (strides in elements; system-major: sys_stride = M, elem_stride = 1; element-major: sys_stride = 1,
 * elem_stride = batch).  a[0] and c[M-1] are not used.  work: device scratch of
 * qz_tridiag_workspace_doubles(M, batch) doubles for the c' of every system; the call allocates nothing. */
size_t qz_tridiag_workspace_doubles(int M, int batch);
int qz_tridiag_thomas_batched(const double* a, const double* b, const double* c, const double* d, double* x, int M,
                              int batch, long long sys_stride, long long elem_stride, double* work, void* stream);

/* Beam deposit: adds the sources of one beam to bsrc (nloc, N+1, nc)
 * (Beam::deposit, src/Beam.cpp:725-785).  Particles are summed per node in index order.
 * work: device scratch of qz_beam_deposit_workspace_bytes(np, nloc, N) bytes. */
size_t qz_beam_deposit_workspace_bytes(int np, int nloc, int N);
int qz_beam_deposit(const qz_radial_grid* grid, const qz_beam_grid* box, int modes, double q, int np,
                    const qz_beam_soa* p, double* bsrc, int nc, void* work, void* stream);
/* The deposits of nbeams beams (charges q[b], np[b] particles, arrays p[b]) in one pass: the same bsrc as
 * qz_beam_deposit of beam 0, then beam 1, ... (Simulation::step_fields' loop over beams,
 * src/Simulation.cpp:642-643), bit for bit.  work: qz_beam_deposit_workspace_bytes(sum of np, nloc, N) bytes.
 * pstep (optional, a host array of nbeams device pointers, as qz_beam_advance_slices marks them): particles
 * whose marker is `marker` or later are left out.  A step run range by range deposits a range after it has pushed
 * earlier ranges; the particles it pushed there were deposited before their push and must not be seen again
 * at their new places (a slow particle can slip into a later range).  live_count (optional, device): adds the
 * live particles whose lower slice is in rows [k0, k0 + nloc) (the first window of the box also those before it,
 * the last one those after it), so the windows of one step count every live particle once. */
int qz_beam_deposit_beams(const qz_radial_grid* grid, const qz_beam_grid* box, int modes, int nbeams, const double* q,
                          const int* np, const qz_beam_soa* p, double* bsrc, int nc, void* work,
                          const int* const* pstep, int marker, unsigned long long* live_count, void* stream);

/* Beam push with the stored fields fld (nloc, N+1, nc): gather, pusher (0 vay, 1 boris, 2 hc, 3 imp),
 * drift; move = 0 is the explicit leapfrog start-up kick (Beam::advance, src/Beam.cpp:948-1131). */
int qz_beam_advance(const qz_radial_grid* grid, const qz_beam_grid* box, int modes, double qm, int pusher, int i0,
                    int np, const qz_beam_soa* p, const double* fld, int nc, double dt_kick, double dt, int move,
                    void* stream);
/* The same push restricted to the particles whose gather slice (after clamping to the box) is in
 * [s_lo, s_hi) and whose marker pstep[i] (np ints, caller-owned) is below `step`; pushed particles get
 * pstep[i] = step.  Markers grow with the step, so a particle a later step has already pushed is not pushed
 * again.  A step pushed range by range, in increasing slice order, then pushes every live particle exactly
 * once, as Beam::advance does (particles only slip to later slices); pushed_count (optional, device) adds the
 * live particles this call pushes.  Not in the reference:
 * it serves the pipelined run, where slice ranges are pushed as soon as their fields exist. */
int qz_beam_advance_slices(const qz_radial_grid* grid, const qz_beam_grid* box, int modes, double qm, int pusher,
                           int np, const qz_beam_soa* p, const double* fld, int nc, double dt_kick, double dt,
                           int move, int s_lo, int s_hi, int step, int* pstep, unsigned long long* pushed_count,
                           void* stream);

/* The 17 weighted beam sums relative to gref and xiref (Beam::sums, src/Beam.cpp:1133-1155): the particles
 * are copied to the host and summed in index order, as the reference does (bit-identical). */
int qz_beam_sums(int np, const qz_beam_soa* p, double gref, double xiref, double* out17_host, void* stream);

/* The same 17 sums reduced on the device into out17 (device memory), without copying the particles:
 * a fixed-shape two-level reduction, reproducible run to run, equal to qz_beam_sums up to the rounding
 * of the summation order.  work: qz_beam_sums_workspace_bytes(np) bytes of device memory.  Replaces the
 * host copy of Beam::sums for large beams (Beam::sums, src/Beam.cpp:1133-1155). */
size_t qz_beam_sums_workspace_bytes(int np);
int qz_beam_sums_device(int np, const qz_beam_soa* p, double gref, double xiref, double* out17, void* work,
                        void* stream);

/* One quasi-static sweep over all slices of one run: per slice the mode deposit, the psi / E_z / B_z and
 * B+ solves, the field store and the Adams-Bashforth push, in one persistent kernel
 * (Simulation::step_fields slice loop, src/Simulation.cpp:727-783).  args points to host memory; the
 * argument blocks are staged in a device buffer kept per (device, stream), allocated on a stream's first
 * call (so that call is not capturable in a graph). */
size_t qz_sweep_workspace_doubles(int M);
/* Workspace of a run swept by `groups` cooperating workgroups (qz_sweep_args.groups > 1): M nodes, np mobile plasma
 * particles (all species), modes 0 or 1.  The workgroups must be resident together: a launch of more workgroups
 * than the device holds at once (one per compute unit) is refused with QZ_ERANGE, and a grid barrier that is not
 * reached within two seconds sets the run's error flag (mws word 2) and ends the launch. */
size_t qz_sweep_multi_workspace_doubles(int M, int np, int groups, int modes);
int qz_sweep(const qz_sweep_args* args, void* stream);

/* Many independent runs in one persistent launch, one workgroup per run, or each its qz_sweep_args.groups
 * cooperating workgroups when any run has groups > 1 (not in the reference). */
int qz_sweep_batched(const qz_sweep_args* runs, int nruns, void* stream);
/* No counterpart in the reference: qz_sweep for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation).  It allocates nothing: the argument block
 * and the run's workspace (args->ws is ignored) live in a static device array of the library, about 0.9 MB per
 * GPU for the life of the process, which limits the grid to 2049 nodes (larger grids: QZ_ERANGE).  Every call on
 * one GPU shares that array: calls whose sweeps overlap in time race, and a launch that continues a slice range
 * (k_begin > 0) reads the B+ warm start the previous call on that GPU left there.  The block is written by a small
 * kernel, so a captured call keeps its arguments in the graph. */
int qz_sweep_noalloc(const qz_sweep_args* args, void* stream);

/* sizeof(qz_sweep_args), for bindings that mirror the struct */
size_t qz_sweep_args_bytes(void);

/* The state one slice block hands to the next (Simulation::make_message, src/Simulation.cpp:851-870;
 * PlasmaSpecies::pack_state / unpack_state, src/PlasmaSpecies.cpp:317-364), packed in device memory: the header
 * [step][dt][last], one zero particle count per beam, per mobile species [Np][x][y][px][py][Delta][w][fresh]
 * [AB history slabs (order, 5)][pushes][xi][xi of each history slot], B+ (r0 i0 r1 i1 r2 i2, M each) and the
 * beam sources of the block's last slice (M x 15, component fastest).  Device copies and small kernels only: the
 * message can go to a peer GPU or a checkpoint file without a host round trip.  pack: the run's state after a
 * launch that ended at slice k_next (one workgroup per run).  unpack: into another run's arrays (its species
 * counts in *args updated, args->load cleared); a launch from k_begin = k_next then continues the sweep bit for
 * bit; guard (optional, device, M x 15) receives the beam-source row. */
size_t qz_message_doubles(const qz_sweep_args* args, int nbeams);
int qz_pack_message(const qz_sweep_args* args, int nbeams, int step, double dt, int last, int k_next, double* msg,
                    void* stream);
int qz_unpack_message(qz_sweep_args* args, int nbeams, const double* msg, double* guard, void* stream);

/* The adjoint of one m = 0 step (no counterpart in the reference, which has no adjoint; the forward it
 * differentiates is Simulation::step_fields, src/Simulation.cpp:636-795, and Beam::advance,
 * src/Beam.cpp:948-1131).  One call walks the slices [k_lo, k_hi) of a one-workgroup run backwards: from dJ/dfld
 * (K, M, 6) it accumulates dJ/dbsrc (K, M, 5), through the plasma's response, slice by slice.  traj holds the
 * plasma state at the start of each slice of the range (a qz_sweep launch over [k_lo, k_hi) with args.traj set:
 * the checkpointed recomputation); ring (order x 5 x np) carries the plasma state's adjoint from one call to the
 * next (zero before the call with k_hi = K); lam_bsrc is accumulated (zero it before the first call).  Particle
 * removals and wall reflections are fixed events.  Grids up to 1023 cells; one workgroup, 512 lanes. */
typedef struct qz_adjoint_args {
    const double* traj;
    int k_lo, k_hi;
    const double* lam_fld;
    double* lam_bsrc;
    double* ring;
    double* work;              /* qz_sweep_adjoint_workspace_doubles(M, np) */
    unsigned long long* ticks; /* optional: += wall-clock ticks of the reverse sweep's phases (QZ_ADJ_NPHASE) */
} qz_adjoint_args;
#define QZ_ADJ_NPHASE 8        /* fields, chi + forces, B+ solves, B+ nodes, deposit_S, Ez/Bz/psi solves, their nodes,
                                  sources + deposit */
size_t qz_sweep_adjoint_workspace_doubles(int M, int np);
int qz_sweep_adjoint(const qz_sweep_args* args, const qz_adjoint_args* adj, void* stream);
/* Adjoints of the beam side of the step (m = 0).  qz_beam_spread_adjoint: dJ/dgamma and the direct dJ/dw of
 * J = sigma / mu, the weighted energy spread of the beam's live particles (Beam::sums / from_sums,
 * src/Beam.cpp:1133-1192), given mu, sigma and the weight sum W.  qz_beam_kick_adjoint: lam_fld += dJ/dfld through
 * the kick of Beam::advance (Vay pusher; started = 0 adds the first push's half back-kick) from dJ/dgamma after it,
 * with the beam as it was before the push.  qz_beam_deposit_adjoint: lam_w += dJ/dw through Beam::deposit
 * (src/Beam.cpp:725-785) from dJ/dbsrc.  qz_profile_gradient: grad = sum_i lam_w_i wmac d profile(u_i) / d v for a
 * piecewise-linear weight profile of nseg values on [u0, u1] (u: each particle's xi as sampled, user units). */
int qz_beam_spread_adjoint(int np, const qz_beam_soa* beam, double mu, double sigma, double wsum, double* lam_gamma,
                           double* lam_w, void* stream);
int qz_beam_kick_adjoint(const qz_radial_grid* grid, const qz_beam_grid* box, double qm, int pusher, int np,
                         const qz_beam_soa* before, const double* fld, int ncf, double dt_kick, int started,
                         const double* lam_gamma, double* lam_fld, void* stream);
int qz_beam_deposit_adjoint(const qz_radial_grid* grid, const qz_beam_grid* box, double q, int np,
                            const qz_beam_soa* before, const double* lam_bsrc, int ncb, double* lam_w, void* stream);
int qz_profile_gradient(int np, const double* u, const double* lam_w, int nseg, double u0, double u1, double wmac,
                        double* grad, void* stream);

/* Frees the argument staging buffer qz_sweep / qz_sweep_batched keep for `stream` on the current device, after
 * the stream's work is done: call it before destroying a stream the sweeps were launched on (a new stream can
 * reuse the handle).  No counterpart in the reference. */
int qz_sweep_release(void* stream);

/* ---- whole simulation (Simulation, src/Simulation.hpp:67-103) on a deck (src/Config.cpp format) ---- */
typedef struct qz_sim qz_sim;
/* The deck reader alone (src/Config.cpp's format; host only, no device work): every key after the overrides as
 * "key = value" lines in key order, written to out when cap exceeds the returned length (bytes, without the
 * terminating 0); a negative code for a deck the reader refuses (qz_last_error() says why). */
long qz_deck_keys(const char* deck, const char* overrides, char* out, long cap);
/* deck: text of an input deck; overrides: "key=value" lines (src/main.cpp:82-93). */
int qz_sim_create(const char* deck, const char* overrides, qz_sim** out);
void qz_sim_destroy(qz_sim* s);
/* info: M K modes NCF NCB NCD nspecies nbeams nsteps next_step */
int qz_sim_info(const qz_sim* s, int* info10);
/* one step of Simulation::run (src/Simulation.cpp:1376-1459): sweep, then the beam push unless last */
int qz_sim_step(qz_sim* s);
int qz_sim_step_fields(qz_sim* s);   /* the sweep only */
/* The part of Simulation::step_fields before the slice loop (beam deposit, plasma load, background; the run's
 * stream is synchronised on return) and the argument block of its sweep, for a caller that launches the sweep
 * itself (qz_sweep, qz_sweep_noalloc) on the run's arrays.  The run's step counter does not advance. */
int qz_sim_prepare_sweep(qz_sim* s, qz_sweep_args* out);
/* One step of n independent runs (same modes) whose sweeps are launched together; beam deposits and pushes run
 * per run.  The runs with one workgroup (their `groups`, see qz_sim_tune_groups and qz_batch_balance) share one
 * persistent launch whose workgroups take them from a queue, longest first; the runs with several workgroups
 * share one multi-workgroup launch beside it, on a second stream.  The workers plus the multi-workgroup runs'
 * workgroups never exceed the compute units.  batch_ms (optional) receives the time from the first launch to
 * the end of both.  Not in the reference: it batches what Simulation::run does for one run. */
int qz_sim_step_batched(qz_sim** sims, int n, double* batch_ms);
int qz_sim_push_beams(qz_sim* s);    /* the beam push of the current step */
/* nrun steps of Simulation::run with up to `depth` steps in flight, each step split into nchunks slice ranges:
 * step n sweeps a range as soon as step n-1 has pushed the beam particles that can reach it.  Bit-identical
 * to nrun calls of qz_sim_step.  Needs a started run (after the first step), the same dt and no output
 * directory (output.dir: the beam log and field files are written between steps).  out2 (optional): particle-slices and beam particles pushed over the nrun steps.  Not in
 * the reference: its MPI ranks pipeline slice blocks instead, with different arithmetic at the block edges. */
int qz_sim_run_pipelined(qz_sim* s, int nrun, int nchunks, int depth, double* out2);
/* what: fld diag bsrc nodes, beam<b>.<x|y|px|py|pz|xi|w>, sp<s>.<x|y|px|py|dl|w>; returns the count */
long qz_sim_get(qz_sim* s, const char* what, double* host_out, long cap);
/* beam log values of beam b (Beam::from_sums of the sums reduced on the device, qz_beam_sums_device order):
 * t npart charge gamma_mean gamma_rms xi_mean xi_rms r_rms emit_nx x_mean y_mean emit_ny sigma_x sigma_y alive */
int qz_sim_beam_diag(qz_sim* s, int b, double* out15);
/* the 17 sums of Beam::sums in the reference's index order (host copy of the particles) and gref, xiref */
int qz_sim_beam_sums(qz_sim* s, int b, double* out17, double* refs2);
/* No counterpart in the reference: the differentiable workload (a witness whose current follows a
 * piecewise-linear profile, deck key <beam>.pwl_weights on a flat-top beam).  set_beam_profile gives beam b
 * new profile values (nseg of them, positive): every macro-particle's weight becomes the flat-top weight
 * times the profile at the xi it was sampled at.  snapshot keeps the beams and the step state as they are;
 * rewind returns to them, so one step can be evaluated again with other profile values. */
int qz_sim_set_beam_profile(qz_sim* s, int b, const double* values, int nseg);
/* One step and the gradient of its objective, by the adjoint (m = 0; no counterpart in the reference): J =
 * gamma_rms / gamma_mean of beam b after the step and dJ/dv of b's nseg weight-profile values.  The sweep runs in
 * checkpointed segments of `every` slices; the reverse sweep recomputes each from its checkpoint (qz_sweep_adjoint).
 * ms2 (optional): milliseconds of the checkpointed sweep and of the reverse part.  The run ends one step on. */
int qz_sim_spread_grad(qz_sim* s, int b, int every, double* J, double* grad, int nseg, double* ms2);
/* Wall-clock ticks (100 MHz clock) of the last qz_sim_spread_grad's reverse-sweep phases, QZ_ADJ_NPHASE doubles. */
int qz_sim_adjoint_ticks(qz_sim* s, double* out);
/* Number of values of beam b's weight profile (<beam>.pwl_weights; 0 without one), or a negative error. */
int qz_sim_profile_size(qz_sim* s, int b);
/* Record chi_0 of every slice of the coming sweeps (qz_sweep_args.chi_out) into a (K, M) device array of the run:
 * the plasma susceptibility that the laser envelope's advance_slice takes (src/Simulation.cpp:735), ready to pass
 * to qz_laser_step.  on = 0 stops and frees it.  qz_sim_get(s, "chi", ...) copies it to the host;
 * qz_sim_chi_device gives its device address (NULL when not recording).  Set before a qz_batch is captured.  In
 * pipelined runs a step sweeps each range after the step before it, so the array ends with the last step's chi_0. */
int qz_sim_record_chi(qz_sim* s, int on);
int qz_sim_chi_device(qz_sim* s, double** ptr);
int qz_sim_snapshot(qz_sim* s);
int qz_sim_rewind(qz_sim* s);
/* times of the last sweep and beam push (ms, HIP events on the simulation's stream) */
int qz_sim_timing(const qz_sim* s, double* out2);
/* work of the last step: particle-slices of the sweep, beam particles pushed, beam particles deposited,
 * sweep kernel ms, beam deposit ms, beam push ms, sweep kernel launches.  After qz_sim_run_pipelined: the sweep
 * kernel ms are the sum of every range launch's time (HIP events around each launch) divided by the steps, and
 * the launches are the ranges of one step; the beam times are 0 (deposits and pushes run inside the ranges). */
int qz_sim_stats(const qz_sim* s, double* out7);
/* wall-clock ticks (100 MHz) of the QZ_NPHASE phases of the last sweep, measured by the first thread of the
 * run (a phase includes that thread's wait at the barrier that ends it): geometry, sort, deposit
 * contributions, deposit node sums, combine, B_z rhs, psi/E_z/B_z solves, W+, S+ gather, S+ contributions,
 * S+ node sums, Picard staging, Picard rhs, Picard solves, store, push */
int qz_sim_phase_ticks(const qz_sim* s, double* out16);
/* The schedule of the last qz_sim_run_pipelined, per range launch (step-major: launch i = step i / nchunks,
 * range i % nchunks), four times in ms from HIP events on the steps' streams, relative to the first range's
 * deposit start: the range's beam deposit starts (its stream's earlier work and the previous step's push it
 * waits for are done), its sweep starts, its sweep ends, its beam push ends.  Writes min(count, cap) quadruples
 * to out (optional) and the ranges per step to *nchunks (optional); returns the count (0 before any pipelined
 * run), negative on error.  Not in the reference. */
long qz_sim_pipeline_timeline(const qz_sim* s, double* out, long cap, int* nchunks);
long qz_sim_lost(const qz_sim* s);
int qz_sim_groups(const qz_sim* s);   /* workgroups per run of the sweep (1: the one-workgroup kernel) */
double qz_sim_time(const qz_sim* s);
int qz_sim_write_fields(qz_sim* s, const char* path);

/* ---- a fixed set of runs stepped together (not in the reference: it batches Simulation::run) ----
 * Every step launches all sweeps together (as qz_sim_step_batched).  With use_graph, once a step
 * repeats the device work of the step before in every run (every beam started, the same dt, a beam push to
 * do, no field output), the whole step -- each run's beam deposit and plasma reload, the batched sweep, each
 * run's beam push -- is captured once as a HIP graph and replayed on later steps.  step_ms receives the
 * time of the batched sweep (eager steps) or of the whole replayed step.  The runs must stay alive and must
 * not be stepped by other calls while the batch exists. */
typedef struct qz_batch qz_batch;
int qz_batch_create(qz_sim** sims, int n, int use_graph, qz_batch** out);
int qz_batch_step(qz_batch* b, double* step_ms);
long qz_batch_replays(const qz_batch* b);   /* steps run as graph replays so far */
void qz_batch_destroy(qz_batch* b);
/* At most max_workers workgroups (0: every compute unit the multi-workgroup runs leave) take the batch's
 * one-workgroup runs from the queue, e.g. to leave compute units to other work.  Results do not depend on it. */
int qz_batch_set_workers(qz_batch* b, int max_workers);
/* Autotuner of the batch's workgroups per run, before the first step.  The runs of one shape (modes, cells,
 * plasma particles) are timed on one of them, the one with the most slices: its first nslices slices with each
 * candidate count that leaves at most 1024 plasma particles per workgroup (the multi-workgroup runs share one
 * launch, which sizes every run's per-workgroup share by the largest), best of two launches.  Each run's sweep is estimated at (time per slice) x (its slices); a
 * plan gives every run whose one-workgroup estimate exceeds a target the fewest workgroups that bring it under
 * the target and list-schedules the other runs, longest first, on the remaining compute units; every estimate
 * is tried as the target and the plan with the smallest estimated sweep is applied to the runs' `groups`.
 * groups_out[n] receives each run's count; est_ms[2] the plan's estimate and the estimate with one workgroup
 * per run (-1 when some run cannot take one).  Results do not depend on the counts. */
int qz_batch_balance(qz_batch* b, const int* candidates, int ncand, int nslices, int* groups_out, double* est_ms);
/* The autotuner's planning step alone, host only: est[i * ncand + c] is run i's estimated sweep (ms) with
 * candidates[c] workgroups (< 0: unusable), ncu the compute units.  Writes groups_out[n] and est_ms[2] as
 * qz_batch_balance does; an error (-1, with a message) when no plan fits the compute units. */
int qz_batch_plan(int n, const int* candidates, int ncand, const double* est, int ncu, int* groups_out,
                  double* est_ms);

/* ---- launch shapes of the beam kernels (not in the reference) ----
 * Threads per block (a multiple of 64 in [64, 1024]) of: the beam push (Beam::advance), the deposit's
 * per-particle kernel (cell keys and records), a slot kept for the bucket sort (the bucket order now comes from
 * a radix sort of fixed shape, so this entry has no effect) and its per-node sums.  One table for the process,
 * 256 each until set.  Results do not depend on the shapes. */
enum { QZ_K_BEAM_ADVANCE = 0, QZ_K_BEAM_PARTICLES = 1, QZ_K_BEAM_SORT = 2, QZ_K_BEAM_NODES = 3, QZ_K_COUNT = 4 };
int qz_set_block(int kernel, int threads);
int qz_get_block(int kernel);
/* Autotuner step: times each of ncand block sizes for `kernel` on scratch copies of the run's largest beam
 * (the push, or the whole deposit into a scratch source array), reps launches each after one warm-up launch;
 * ms_out[i] receives candidate i's mean time, and the fastest is set.  The run's own state is not changed. */
int qz_sim_autotune(qz_sim* s, int kernel, const int* candidates, int ncand, int reps, double* ms_out);
/* Autotuner of the sweep's workgroups per run: each candidate count (1: the one-workgroup kernel, when the run fits
 * it) sweeps the first nslices slices of the coming step, timed with HIP events into ms_out[i] (-1: a count the
 * run cannot use); the fastest is kept for the run.  Before the first step only.  Results do not depend on the
 * count.  Not in the reference. */
int qz_sim_tune_groups(qz_sim* s, const int* candidates, int ncand, int nslices, double* ms_out);
/* "<gfx arch name>|<compute units>CU" of the current device, the key under which launch shapes are cached */
int qz_device_tag(char* out, int n);

/* ---- laser envelope (src/Laser.hpp:47-108, src/Laser.cpp) ----
 * The complex envelope a^ of one box: K slices (slice 0 at the head of the box, internal xi = xi_min + k dxi) of
 * M radial nodes, (K, M, 2) doubles (Re, Im) in device memory.  laser.l = 0, no ionization losses, one box on one
 * GPU (the head of the box is the first slice: no guard rows).  scheme: 0 trapezoid, 1 BDF2 (the reference's
 * default), 2 backward Euler in xi (laser.xi_scheme); circular: 1 for circular polarisation (<a^2> = |a^|^2). */
typedef struct qz_laser qz_laser;
/* Laser::Laser without the pulse (src/Laser.cpp:24-84): the grid of M nodes (nodes[0] = 0, increasing), the box,
 * the time step dt of the beam push (dt <= 0: the envelope is never advanced, src/Laser.cpp:318) and
 * k0 = omega0 / omega_p > 1.  The envelope starts at zero.  Grids up to 1861 nodes (QZ_ERANGE beyond). */
int qz_laser_create(const double* nodes, int M, int K, double xi_min, double dxi, double dt, double k0, int scheme,
                    int circular, qz_laser** out);
void qz_laser_destroy(qz_laser* L);
/* the Gaussian pulse of Laser::Laser (src/Laser.cpp:195-236, laser.l = 0): laser.a0, w0, L0, xi0 (internal xi)
 * and focus (the distance to the focal plane), formed on the host and uploaded */
int qz_laser_gaussian(qz_laser* L, double a0, double w0, double L0, double xi0, double focus);
/* the current envelope from / to host memory: (K, M, 2) doubles */
int qz_laser_set_envelope(qz_laser* L, const double* a_host);
/* what = 0: the current envelope (K, M, 2) (Laser::current); 1: the ponderomotive arrays of the last step
 * (K, M, 3): <a^2>, d<a^2>/dr, d<a^2>/dxi at the level the step started from (Laser::ponderomotive) */
int qz_laser_get(qz_laser* L, int what, double* out_host);
/* the device address of the same arrays (valid until the next qz_laser_step, which exchanges the envelope's
 * two buffers) */
int qz_laser_device(qz_laser* L, int what, double** ptr);
/* One time step (src/Simulation.cpp:731-735 and 784, over every slice): Laser::prepare_slice of every slice, then
 * Laser::advance_slice from the head of the box with the plasma susceptibility chi of each slice, (K, M) doubles
 * in device memory (NULL: vacuum); the advanced envelope becomes the current one (Laser::begin_step).
 * stream: NULL runs on the laser's own stream and waits for it; otherwise the call is asynchronous on `stream`,
 * and the caller synchronizes `stream` before the next qz_laser_* call. */
int qz_laser_step(qz_laser* L, const double* chi, void* stream);
/* The same step for n distinct boxes at once (not in the reference): each box's march is one workgroup, so the
 * boxes march side by side; box i's results are those of qz_laser_step(Ls[i], chis[i], ...).  chis: n device
 * arrays (or NULL entries, or a NULL array: vacuum).  The call first waits for earlier work on its stream (NULL:
 * the first box's own stream, waited for at the end too).  Boxes may differ in grid, box, scheme and dt. */
int qz_laser_step_batched(qz_laser* const* Ls, int n, const double* const* chis, void* stream);
/* Laser::sums (src/Laser.cpp:421-438) of the current envelope, summed on the host in slice and node order:
 * {sum |a|^2 dV, sum |a|^2 xi dV (the user's xi = -xi), sum |a|^2 r^2 dV, max |a|, sum |a|^2 xi^2 dV} */
int qz_laser_sums(qz_laser* L, double* out5);

#ifdef __cplusplus
}
#endif
#endif
