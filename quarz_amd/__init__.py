"""quarz_amd: the quasi-static time step of QUARZ on MI355X (gfx950).

The product is libquarz_amd.so (C++ host + gfx950 HIP kernels behind the C-ABI of include/quarz_amd.h).
This module is a thin ctypes driver over it: `Simulation` mirrors the reference's Simulation on a deck
(src/Simulation.hpp:67-103), and the kernel-level entry points take device pointers (e.g. of torch tensors).
There is no CPU fallback: without the built library every call raises.
"""
import ctypes as C
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libquarz_amd.so")

_lib = None

DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)


class QuarzError(RuntimeError):
    pass


class RadialGridC(C.Structure):
    _fields_ = [("r", C.c_void_p), ("inv_h", C.c_void_p), ("inv_V", C.c_void_p), ("N", C.c_int), ("R", C.c_double)]


class SweepArgsHeadC(C.Structure):
    """The leading fields of qz_sweep_args (include/quarz_amd.h): enough to set a launch's slice range on a block
    that qz_sim_prepare_sweep filled."""
    _fields_ = [("grid", RadialGridC), ("modes", C.c_int), ("picard", C.c_int), ("nslices", C.c_int),
                ("k_begin", C.c_int), ("k_end", C.c_int)]


class BeamGridC(C.Structure):
    _fields_ = [("xi_min", C.c_double), ("dxi", C.c_double), ("nxi", C.c_int), ("k0", C.c_int), ("nloc", C.c_int),
                ("ngp", C.c_int)]


class BeamSoAC(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("x", "y", "px", "py", "pz", "xi", "w")]


def lib():
    """Load libquarz_amd.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise QuarzError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        # torch ships its own HIP runtime.  Loading it first makes this library bind to the same runtime;
        # if this library's runtime starts first, torch's own cannot open the device (torch.cuda.is_available()
        # turns False) and the process has two runtimes.
        import torch  # noqa: F401
        L = C.CDLL(LIB_PATH)
        L.qz_last_error.restype = C.c_char_p
        L.qz_device_info.argtypes = [C.c_char_p, C.c_int, IP]
        L.qz_tridiag_workspace_doubles.restype = C.c_size_t
        L.qz_tridiag_workspace_doubles.argtypes = [C.c_int, C.c_int]
        L.qz_tridiag_thomas_batched.argtypes = [C.c_void_p] * 5 + [C.c_int, C.c_int, C.c_longlong, C.c_longlong,
                                                                  C.c_void_p, C.c_void_p]
        L.qz_beam_deposit_workspace_bytes.restype = C.c_size_t
        L.qz_beam_deposit_workspace_bytes.argtypes = [C.c_int, C.c_int, C.c_int]
        L.qz_beam_deposit.argtypes = [C.POINTER(RadialGridC), C.POINTER(BeamGridC), C.c_int, C.c_double, C.c_int,
                                      C.POINTER(BeamSoAC), C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.qz_beam_deposit_beams.argtypes = [C.POINTER(RadialGridC), C.POINTER(BeamGridC), C.c_int, C.c_int, DP, IP,
                                            C.POINTER(BeamSoAC), C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                            C.c_int, C.c_void_p, C.c_void_p]
        L.qz_beam_advance.argtypes = [C.POINTER(RadialGridC), C.POINTER(BeamGridC), C.c_int, C.c_double, C.c_int,
                                      C.c_int, C.c_int, C.POINTER(BeamSoAC), C.c_void_p, C.c_int, C.c_double,
                                      C.c_double, C.c_int, C.c_void_p]
        L.qz_beam_sums.argtypes = [C.c_int, C.POINTER(BeamSoAC), C.c_double, C.c_double, DP, C.c_void_p]
        L.qz_beam_sums_workspace_bytes.restype = C.c_size_t
        L.qz_beam_sums_workspace_bytes.argtypes = [C.c_int]
        L.qz_beam_sums_device.argtypes = [C.c_int, C.POINTER(BeamSoAC), C.c_double, C.c_double, C.c_void_p,
                                          C.c_void_p, C.c_void_p]
        L.qz_sweep_workspace_doubles.restype = C.c_size_t
        L.qz_sweep_workspace_doubles.argtypes = [C.c_int]
        L.qz_sim_create.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]
        L.qz_sim_destroy.argtypes = [C.c_void_p]
        L.qz_sim_info.argtypes = [C.c_void_p, IP]
        L.qz_deck_keys.restype = C.c_long
        L.qz_deck_keys.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_long]
        L.qz_sim_step.argtypes = [C.c_void_p]
        L.qz_sim_step_fields.argtypes = [C.c_void_p]
        L.qz_sim_prepare_sweep.argtypes = [C.c_void_p, C.c_void_p]
        L.qz_sim_tune_groups.argtypes = [C.c_void_p, IP, C.c_int, C.c_int, DP]
        L.qz_sweep.argtypes = [C.c_void_p, C.c_void_p]
        L.qz_sweep_noalloc.argtypes = [C.c_void_p, C.c_void_p]
        L.qz_sweep_args_bytes.restype = C.c_size_t
        L.qz_sweep_args_bytes.argtypes = []
        L.qz_message_doubles.restype = C.c_size_t
        L.qz_message_doubles.argtypes = [C.c_void_p, C.c_int]
        L.qz_pack_message.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_void_p,
                                      C.c_void_p]
        L.qz_unpack_message.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.qz_sim_step_batched.argtypes = [C.POINTER(C.c_void_p), C.c_int, DP]
        L.qz_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        L.qz_batch_step.argtypes = [C.c_void_p, DP]
        L.qz_batch_balance.argtypes = [C.c_void_p, IP, C.c_int, C.c_int, IP, DP]
        L.qz_batch_set_workers.argtypes = [C.c_void_p, C.c_int]
        L.qz_batch_plan.argtypes = [C.c_int, IP, C.c_int, DP, C.c_int, IP, DP]
        L.qz_batch_replays.argtypes = [C.c_void_p]
        L.qz_batch_replays.restype = C.c_long
        L.qz_batch_destroy.argtypes = [C.c_void_p]
        L.qz_batch_destroy.restype = None
        L.qz_set_block.argtypes = [C.c_int, C.c_int]
        L.qz_get_block.argtypes = [C.c_int]
        L.qz_sim_autotune.argtypes = [C.c_void_p, C.c_int, IP, C.c_int, C.c_int, DP]
        L.qz_device_tag.argtypes = [C.c_char_p, C.c_int]
        L.qz_sim_run_pipelined.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, DP]
        L.qz_sim_push_beams.argtypes = [C.c_void_p]
        L.qz_sim_get.restype = C.c_long
        L.qz_sim_get.argtypes = [C.c_void_p, C.c_char_p, DP, C.c_long]
        L.qz_sim_beam_diag.argtypes = [C.c_void_p, C.c_int, DP]
        L.qz_sim_beam_sums.argtypes = [C.c_void_p, C.c_int, DP, DP]
        L.qz_sim_timing.argtypes = [C.c_void_p, DP]
        L.qz_sim_stats.argtypes = [C.c_void_p, DP]
        L.qz_sim_phase_ticks.argtypes = [C.c_void_p, DP]
        L.qz_sim_pipeline_timeline.argtypes = [C.c_void_p, DP, C.c_long, C.POINTER(C.c_int)]
        L.qz_sim_pipeline_timeline.restype = C.c_long
        L.qz_sim_groups.argtypes = [C.c_void_p]
        L.qz_sim_lost.restype = C.c_long
        L.qz_sim_lost.argtypes = [C.c_void_p]
        L.qz_sim_time.restype = C.c_double
        L.qz_sim_time.argtypes = [C.c_void_p]
        L.qz_sim_write_fields.argtypes = [C.c_void_p, C.c_char_p]
        L.qz_sim_set_beam_profile.argtypes = [C.c_void_p, C.c_int, DP, C.c_int]
        L.qz_sim_snapshot.argtypes = [C.c_void_p]
        L.qz_sim_spread_grad.argtypes = [C.c_void_p, C.c_int, C.c_int, DP, DP, C.c_int, DP]
        L.qz_sim_profile_size.argtypes = [C.c_void_p, C.c_int]
        L.qz_sim_adjoint_ticks.argtypes = [C.c_void_p, DP]
        L.qz_sweep_adjoint_workspace_doubles.restype = C.c_size_t
        L.qz_laser_create.argtypes = [DP, C.c_int, C.c_int] + [C.c_double] * 4 + [C.c_int, C.c_int,
                                                                                 C.POINTER(C.c_void_p)]
        L.qz_laser_destroy.argtypes = [C.c_void_p]
        L.qz_laser_destroy.restype = None
        L.qz_laser_gaussian.argtypes = [C.c_void_p] + [C.c_double] * 5
        L.qz_laser_set_envelope.argtypes = [C.c_void_p, DP]
        L.qz_laser_get.argtypes = [C.c_void_p, C.c_int, DP]
        L.qz_laser_device.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]
        L.qz_laser_step.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.qz_laser_sums.argtypes = [C.c_void_p, DP]
        L.qz_sim_record_chi.argtypes = [C.c_void_p, C.c_int]
        L.qz_sim_chi_device.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        L.qz_laser_step_batched.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p), C.c_void_p]
        L.qz_sweep_adjoint_workspace_doubles.argtypes = [C.c_int, C.c_int]
        L.qz_sweep_adjoint.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.qz_beam_spread_adjoint.argtypes = [C.c_int, C.POINTER(BeamSoAC), C.c_double, C.c_double, C.c_double,
                                             C.c_void_p, C.c_void_p, C.c_void_p]
        L.qz_beam_kick_adjoint.argtypes = [C.POINTER(RadialGridC), C.POINTER(BeamGridC), C.c_double, C.c_int, C.c_int,
                                           C.POINTER(BeamSoAC), C.c_void_p, C.c_int, C.c_double, C.c_int,
                                           C.c_void_p, C.c_void_p, C.c_void_p]
        L.qz_beam_deposit_adjoint.argtypes = [C.POINTER(RadialGridC), C.POINTER(BeamGridC), C.c_double, C.c_int,
                                              C.POINTER(BeamSoAC), C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.qz_profile_gradient.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_double,
                                          C.c_double, C.c_void_p, C.c_void_p]
        L.qz_sim_rewind.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def last_error():
    return lib().qz_last_error().decode()


def check(rc):
    if rc != 0:
        raise QuarzError(last_error())
    return rc


def device_info():
    buf = C.create_string_buffer(64)
    ncu = C.c_int(0)
    check(lib().qz_device_info(buf, 64, C.byref(ncu)))
    return buf.value.decode(), ncu.value


BEAM_ARRAYS = ("x", "y", "px", "py", "pz", "xi", "w")
DIAG_KEYS = ["t", "npart", "charge", "gamma_mean", "gamma_rms", "xi_mean", "xi_rms", "r_rms", "emit_nx", "x_mean",
             "y_mean", "emit_ny", "sigma_x", "sigma_y", "alive"]


def deck_keys(deck, overrides=None):
    """The keys of a deck (text) after key=value overrides, as read by the library's deck reader: {key: value}."""
    ov = "\n".join(f"{k}={v}" for k, v in (overrides or {}).items())
    n = lib().qz_deck_keys(deck.encode(), ov.encode(), None, 0)
    if n < 0:
        raise QuarzError(last_error())
    buf = C.create_string_buffer(n + 1)
    lib().qz_deck_keys(deck.encode(), ov.encode(), buf, n + 1)
    out = {}
    for line in buf.value.decode().splitlines():
        k, _, v = line.partition(" = ")
        out[k] = v
    return out


class Simulation:
    """One QUARZ run on a deck (text or path) with key=value overrides; the step runs on the GPU."""

    def __init__(self, deck, overrides=None):
        if os.path.exists(deck):
            deck = open(deck).read()
        ov = "\n".join(f"{k}={v}" for k, v in (overrides or {}).items())
        h = C.c_void_p()
        check(lib().qz_sim_create(deck.encode(), ov.encode(), C.byref(h)))
        self.h = h
        info = (C.c_int * 10)()
        check(lib().qz_sim_info(self.h, info))
        self.M, self.K, self.m1, self.NCF, self.NCB, self.NCD, self.nspecies, self.nbeams, self.nsteps = info[:9]

    def close(self):
        if getattr(self, "h", None):
            lib().qz_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:   # interpreter shutdown: the library may already be gone
            pass

    def step(self):
        check(lib().qz_sim_step(self.h))

    def step_fields(self):
        check(lib().qz_sim_step_fields(self.h))

    def push_beams(self):
        check(lib().qz_sim_push_beams(self.h))

    def timing(self):
        """(sweep ms, beam push ms) of the last step, from HIP events on the run's stream."""
        out = np.zeros(2)
        check(lib().qz_sim_timing(self.h, out.ctypes.data_as(DP)))
        return out

    def stats(self):
        """Work of the last step: particle_slices, beam_pushed, beam_deposited, sweep_ms, beam_dep_ms, beam_push_ms,
        sweep_launches (after run_pipelined: sweep_ms sums every range launch of a step, over the steps)."""
        out = np.zeros(7)
        check(lib().qz_sim_stats(self.h, out.ctypes.data_as(DP)))
        return dict(zip(["particle_slices", "beam_pushed", "beam_deposited", "sweep_ms", "beam_dep_ms", "beam_push_ms",
                         "sweep_launches"], out))

    PHASES = ["geometry", "sort", "dep_contrib", "dep_sums", "combine", "bz_rhs", "solves", "wplus", "s_gather",
              "s_contrib", "s_sums", "picard_stage", "picard_rhs", "picard_solve", "store", "push"]

    def autotune(self, candidates=(64, 128, 256, 512, 1024), reps=10, cache=None):
        """Choose the beam kernels' threads per block for this run's shapes: one kernel at a time (the others
        at their current choice), every candidate timed on scratch copies of the largest beam.  The choice
        is applied to the process (set_block) and returned as {kernel: threads}.  With `cache` (a JSON file),
        a stored choice for the same device and shapes is applied without timing, and new ones are stored."""
        sizes = sorted(self.get(f"beam{b}.w").size for b in range(self.nbeams))
        key = f"{_device_name()}|m{int(self.m1)}|M{self.M}|K{self.K}|nb{sizes[-1] if sizes else 0}"
        table = {}
        if cache and os.path.exists(cache):
            with open(cache) as f:
                table = json.load(f)
        if key in table:
            for name, threads in table[key].items():
                set_block(name, threads)
            return dict(table[key])
        cand = (C.c_int * len(candidates))(*candidates)
        out = np.zeros(len(candidates))
        chosen, times = {}, {}
        for k, name in enumerate(KERNELS):
            check(lib().qz_sim_autotune(self.h, k, cand, len(candidates), reps, out.ctypes.data_as(DP)))
            chosen[name] = get_block(name)
            times[name] = {int(c): float(t) for c, t in zip(candidates, out)}
        self.autotune_ms = times
        if cache:
            table[key] = chosen
            os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
            with open(cache, "w") as f:
                json.dump(table, f, indent=1)
        return chosen

    def tune_groups(self, candidates, nslices=64):
        """Workgroups per run for this run's sweep: each candidate count sweeps the first `nslices` slices of the
        coming step (before the first step); the fastest is kept.  Returns {count: ms} (-1: not usable)."""
        cand = (C.c_int * len(candidates))(*candidates)
        out = np.zeros(len(candidates))
        check(lib().qz_sim_tune_groups(self.h, cand, len(candidates), int(nslices), out.ctypes.data_as(DP)))
        return {int(c): float(t) for c, t in zip(candidates, out)}

    def run_pipelined(self, steps, chunks=64, depth=8):
        """`steps` steps with up to `depth` in flight, each split into `chunks` slice ranges (bit-identical to
        calling step() `steps` times; the run must have taken its first step).  Returns (particle-slices,
        beam particles pushed) over the steps."""
        out = np.zeros(2)
        check(lib().qz_sim_run_pipelined(self.h, int(steps), int(chunks), int(depth), out.ctypes.data_as(DP)))
        return float(out[0]), float(out[1])

    def pipeline_timeline(self):
        """The last run_pipelined's schedule: an array (steps, ranges, 4) of ms per range launch -- deposit start,
        sweep start, sweep end, push end -- relative to the first range's deposit start."""
        nc = C.c_int(0)
        n = lib().qz_sim_pipeline_timeline(self.h, None, 0, C.byref(nc))
        check(int(min(n, 0)))
        out = np.zeros(4 * n)
        lib().qz_sim_pipeline_timeline(self.h, out.ctypes.data_as(DP), n, None)
        return out.reshape(-1, max(nc.value, 1), 4) if n else out.reshape(0, 0, 4)

    def phase_ms(self):
        """Milliseconds of each sweep phase of the last sweep (one thread's wall clock, 100 MHz ticks)."""
        out = np.zeros(len(self.PHASES))
        check(lib().qz_sim_phase_ticks(self.h, out.ctypes.data_as(DP)))
        return dict(zip(self.PHASES, out * 1e-5))

    def lost(self):
        return lib().qz_sim_lost(self.h)

    @property
    def groups(self):
        """Workgroups of this run's sweep (1: the one-workgroup kernel)."""
        return lib().qz_sim_groups(self.h)

    @property
    def t(self):
        return lib().qz_sim_time(self.h)

    def get(self, what):
        n = lib().qz_sim_get(self.h, what.encode(), None, 0)
        if n < 0:
            raise KeyError(f"{what}: {last_error()}")
        out = np.zeros(n)
        lib().qz_sim_get(self.h, what.encode(), out.ctypes.data_as(DP), n)
        return out

    @property
    def fld(self):
        return self.get("fld").reshape(self.K, self.M, self.NCF)

    def record_chi(self, on=True):
        """Record chi_0 of every slice of the coming sweeps (the plasma susceptibility a Laser step takes)."""
        check(lib().qz_sim_record_chi(self.h, int(bool(on))))

    def chi(self):
        """The recorded chi_0, (K, M), host copy."""
        return self.get("chi").reshape(self.K, self.M)

    def chi_device(self):
        """The device address of the recorded chi_0 (K, M), for Laser.step; 0 when not recording."""
        p = C.c_void_p()
        check(lib().qz_sim_chi_device(self.h, C.byref(p)))
        return p.value or 0

    @property
    def diag(self):
        return self.get("diag").reshape(self.K, self.M, self.NCD)

    @property
    def bsrc(self):
        return self.get("bsrc").reshape(self.K, self.M, self.NCB)

    @property
    def nodes(self):
        return self.get("nodes")

    def beam(self, b):
        return {k: self.get(f"beam{b}.{k}") for k in BEAM_ARRAYS}

    def beam_sums(self, b):
        out, refs = np.zeros(17), np.zeros(2)
        check(lib().qz_sim_beam_sums(self.h, b, out.ctypes.data_as(DP), refs.ctypes.data_as(DP)))
        return out, refs

    def beam_diag(self, b):
        out = np.zeros(15)
        check(lib().qz_sim_beam_diag(self.h, b, out.ctypes.data_as(DP)))
        return dict(zip(DIAG_KEYS, out))

    def write_fields(self, path):
        check(lib().qz_sim_write_fields(self.h, path.encode()))

    def set_beam_profile(self, b, values):
        """New values of beam b's piecewise-linear current profile (deck key <beam>.pwl_weights): every
        macro-particle's weight becomes the flat-top weight times the profile at its sampled xi."""
        v = np.ascontiguousarray(values, dtype=np.float64)
        check(lib().qz_sim_set_beam_profile(self.h, b, v.ctypes.data_as(DP), v.size))

    def spread_grad(self, b, every=64):
        """One step, and by the adjoint: J = gamma_rms / gamma_mean of beam b after it and dJ/dv of b's weight
        profile (m = 0; the sweep checkpointed every `every` slices).  Returns (J, grad, (forward ms, reverse ms))."""
        nseg = int(self.profile_size(b))
        J = np.zeros(1)
        g = np.zeros(nseg)
        ms = np.zeros(2)
        check(lib().qz_sim_spread_grad(self.h, b, every, J.ctypes.data_as(DP), g.ctypes.data_as(DP), nseg,
                                       ms.ctypes.data_as(DP)))
        return J[0], g, (ms[0], ms[1])

    ADJ_PHASES = ["fields", "chi_forces", "bplus_solves", "bplus_nodes", "deposit_S", "poisson_solves",
                  "poisson_nodes", "sources_deposit"]

    def adjoint_ticks(self):
        """Wall-clock ticks (100 MHz) of the last spread_grad's reverse-sweep phases."""
        out = np.zeros(len(self.ADJ_PHASES))
        check(lib().qz_sim_adjoint_ticks(self.h, out.ctypes.data_as(DP)))
        return dict(zip(self.ADJ_PHASES, out))

    def profile_size(self, b):
        """Number of values of beam b's weight profile (0 without one)."""
        n = lib().qz_sim_profile_size(self.h, b)
        if n < 0:
            raise QuarzError(last_error())
        return n

    def snapshot(self):
        """Keep the beams and the step state as they are now (see rewind)."""
        check(lib().qz_sim_snapshot(self.h))

    def rewind(self):
        """Return the beams and the step state to the last snapshot."""
        check(lib().qz_sim_rewind(self.h))


def step_batched(sims):
    """One step of every Simulation in `sims` (same modes), their sweeps in one persistent launch.
    Returns the batched sweep time in ms."""
    arr = (C.c_void_p * len(sims))(*[s.h.value for s in sims])
    ms = C.c_double(0)
    check(lib().qz_sim_step_batched(arr, len(sims), C.byref(ms)))
    return ms.value


# beam kernels with a tunable launch shape (QZ_K_* of quarz_amd.h, in that order)
KERNELS = ("beam_advance", "beam_particles", "beam_sort", "beam_nodes")


def set_block(kernel, threads):
    """Threads per block of a beam kernel (name in KERNELS): a multiple of 64 in [64, 1024], process-wide."""
    check(lib().qz_set_block(KERNELS.index(kernel), int(threads)))


def get_block(kernel):
    return lib().qz_get_block(KERNELS.index(kernel))


def _device_name():
    buf = C.create_string_buffer(256)
    check(lib().qz_device_tag(buf, 256))
    return buf.value.decode()


class Batch:
    """A fixed set of Simulations (same modes) stepped together: the sweeps of one-workgroup runs in one
    persistent launch that takes them from a queue, longest first, those of multi-workgroup runs in one launch
    beside it, and with graph=True the steady-state step (beam deposits, plasma reloads, the sweep launches,
    beam pushes) replayed as one captured HIP graph.  The Simulations must outlive the Batch and must not be
    stepped otherwise."""

    def __init__(self, sims, graph=True, workers=0):
        self.sims = list(sims)
        arr = (C.c_void_p * len(self.sims))(*[s.h.value for s in self.sims])
        self.h = C.c_void_p()
        check(lib().qz_batch_create(arr, len(self.sims), 1 if graph else 0, C.byref(self.h)))
        if workers:
            self.set_workers(workers)

    def set_workers(self, n):
        """At most n workgroups take the one-workgroup runs from the queue (0: no cap)."""
        check(lib().qz_batch_set_workers(self.h, int(n)))

    def step(self):
        """One step of every run; returns the batched sweep time (eager steps) or the whole step's time
        (graph replays) in ms."""
        ms = C.c_double(0)
        check(lib().qz_batch_step(self.h, C.byref(ms)))
        return ms.value

    def balance(self, candidates=(1, 2, 3, 4, 6, 8), nslices=256):
        """Workgroups per run, before the first step: every shape of run (modes, cells, plasma particles) is
        timed over its first `nslices` slices with each candidate count, each run's sweep is estimated from its
        slices, and the plan with the smallest estimated batched sweep is applied (qz_batch_balance).  Returns
        {"groups": [per run], "est_ms": plan estimate, "est_one_ms": estimate with one workgroup per run}."""
        cand = (C.c_int * len(candidates))(*candidates)
        g = (C.c_int * len(self.sims))()
        est = np.zeros(2)
        check(lib().qz_batch_balance(self.h, cand, len(candidates), int(nslices), g, est.ctypes.data_as(DP)))
        return {"groups": list(g), "est_ms": float(est[0]), "est_one_ms": float(est[1])}

    @property
    def replays(self):
        return lib().qz_batch_replays(self.h)

    def close(self):
        if self.h and self.h.value:
            lib().qz_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_plan(est_ms, candidates, ncu=256):
    """The batch autotuner's planning step on given estimates (host only): est_ms[i][c] is run i's sweep time with
    candidates[c] workgroups (< 0: unusable).  Returns {"groups", "est_ms", "est_one_ms"} as Batch.balance."""
    est = np.ascontiguousarray(est_ms, dtype=np.float64)
    n, ncand = est.shape
    assert ncand == len(candidates)
    cand = (C.c_int * ncand)(*candidates)
    g = (C.c_int * n)()
    out = np.zeros(2)
    check(lib().qz_batch_plan(n, cand, ncand, est.ctypes.data_as(DP), int(ncu), g, out.ctypes.data_as(DP)))
    return {"groups": list(g), "est_ms": float(out[0]), "est_one_ms": float(out[1])}


class Laser:
    """The laser envelope of one box (src/Laser.hpp:47-108) on the GPU: K slices of the radial grid `nodes`, slice 0
    at the head of the box (internal xi = xi_min + k dxi).  scheme: "trapezoid", "bdf2" (the reference's default)
    or "euler" in xi; laser.l = 0, no ionization losses.  step(chi) advances the envelope by one time step dt with
    the plasma susceptibility chi (K, M) of every slice (None: vacuum)."""

    SCHEMES = {"trapezoid": 0, "bdf2": 1, "euler": 2}

    def __init__(self, nodes, K, xi_min, dxi, dt, k0, scheme="bdf2", circular=False):
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        self.M, self.K = len(self.nodes), int(K)
        self.xi_min, self.dxi = float(xi_min), float(dxi)
        h = C.c_void_p()
        check(lib().qz_laser_create(self.nodes.ctypes.data_as(DP), self.M, self.K, self.xi_min, self.dxi, float(dt),
                                    float(k0), self.SCHEMES[scheme], int(bool(circular)), C.byref(h)))
        self.h = h

    def gaussian(self, a0, w0, L0, xi0, focus=0.0):
        """The Gaussian pulse of Laser::Laser (xi0 internal)."""
        check(lib().qz_laser_gaussian(self.h, a0, w0, L0, xi0, focus))

    def set_envelope(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == (self.K, self.M, 2)
        check(lib().qz_laser_set_envelope(self.h, a.ctypes.data_as(DP)))

    def envelope(self):
        out = np.zeros((self.K, self.M, 2))
        check(lib().qz_laser_get(self.h, 0, out.ctypes.data_as(DP)))
        return out

    def ponderomotive(self):
        """(K, M, 3) of the last step: <a^2>, d<a^2>/dr, d<a^2>/dxi at the level the step started from."""
        out = np.zeros((self.K, self.M, 3))
        check(lib().qz_laser_get(self.h, 1, out.ctypes.data_as(DP)))
        return out

    def step(self, chi=None, stream=None):
        """One time step; chi: a float64 device tensor (K, M), the device address of such an array (for example
        Simulation.chi_device() of a run on the same grid and box), or None (vacuum)."""
        if isinstance(chi, int):
            ptr = C.c_void_p(chi) if chi else None
        elif chi is not None:
            assert chi.dtype.itemsize == 8 and tuple(chi.shape) == (self.K, self.M) and chi.is_contiguous()
            ptr = C.c_void_p(chi.data_ptr())
        else:
            ptr = None
        check(lib().qz_laser_step(self.h, ptr, stream))

    def sums(self):
        out = np.zeros(5)
        check(lib().qz_laser_sums(self.h, out.ctypes.data_as(DP)))
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().qz_laser_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def laser_step_batched(lasers, chis=None, stream=None):
    """One step of several distinct Laser boxes in one launch per kernel (each box's march is one workgroup);
    chis: None or one float64 device tensor (K, M) or None per box."""
    n = len(lasers)
    hs = (C.c_void_p * n)(*[x.h.value for x in lasers])
    cs = None
    if chis is not None:
        assert len(chis) == n
        for x, c in zip(lasers, chis):
            if c is not None:
                assert c.dtype.itemsize == 8 and tuple(c.shape) == (x.K, x.M) and c.is_contiguous()
        cs = (C.c_void_p * n)(*[c.data_ptr() if c is not None else None for c in chis])
    check(lib().qz_laser_step_batched(hs, n, cs, stream))


# ---------------------------------------------------------------------------- kernel-level entry points
def _ptr(t):
    return C.c_void_p(t.data_ptr())


def thomas_batched(a, b, c, d, x, work=None, element_major=False, stream=None):
    """Batched Thomas solve on float64 device tensors: (batch, M) system-major, or (M, batch) with
    element_major=True (element i of system s at i * batch + s).  work: a device tensor of at least
    qz_tridiag_workspace_doubles(M, batch) doubles (made here if None)."""
    if element_major:
        M, batch = d.shape
        ss, es = 1, batch
    else:
        batch, M = d.shape
        ss, es = M, 1
    if work is None:
        import torch
        work = torch.empty(max(1, lib().qz_tridiag_workspace_doubles(M, batch)), dtype=torch.float64, device=d.device)
    check(lib().qz_tridiag_thomas_batched(_ptr(a), _ptr(b), _ptr(c), _ptr(d), _ptr(x), M, batch, ss, es, _ptr(work),
                                          stream))


def radial_grid(r, inv_h, inv_V):
    """qz_radial_grid over device tensors (r: N+1, inv_h: N, inv_V: N+1)."""
    return RadialGridC(r.data_ptr(), inv_h.data_ptr(), inv_V.data_ptr(), r.numel() - 1, float(r[-1].item()))


def beam_soa(p):
    return BeamSoAC(*[p[k].data_ptr() for k in BEAM_ARRAYS])


def beam_deposit(grid, box, modes, q, p, bsrc, work, stream=None):
    soa = beam_soa(p)
    nc = bsrc.shape[-1]
    check(lib().qz_beam_deposit(C.byref(grid), C.byref(box), modes, q, p["x"].numel(), C.byref(soa), _ptr(bsrc), nc,
                                _ptr(work), stream))


def beam_deposit_beams(grid, box, modes, qs, ps, bsrc, work, stream=None):
    """Deposits of several beams (charges qs, particle dicts ps) in one pass, as one deposit after another."""
    n = len(ps)
    soas = (BeamSoAC * n)(*[beam_soa(p) for p in ps])
    q = (C.c_double * n)(*qs)
    npa = (C.c_int * n)(*[p["x"].numel() for p in ps])
    check(lib().qz_beam_deposit_beams(C.byref(grid), C.byref(box), modes, n, q, npa, soas, _ptr(bsrc), bsrc.shape[-1],
                                      _ptr(work), None, 0, None, stream))


def beam_deposit_workspace_bytes(np_, nloc, N):
    return lib().qz_beam_deposit_workspace_bytes(np_, nloc, N)


def beam_advance(grid, box, modes, qm, pusher, p, fld, dt_kick, dt, move, i0=0, stream=None):
    soa = beam_soa(p)
    nc = fld.shape[-1]
    check(lib().qz_beam_advance(C.byref(grid), C.byref(box), modes, qm, pusher, i0, p["x"].numel(), C.byref(soa),
                                _ptr(fld), nc, dt_kick, dt, int(move), stream))


def beam_sums(p, gref, xiref, stream=None):
    """Beam::sums of device SoA tensors, summed on the host in index order (bit-identical to the reference)."""
    soa = beam_soa(p)
    out = np.zeros(17)
    check(lib().qz_beam_sums(p["x"].numel(), C.byref(soa), gref, xiref, out.ctypes.data_as(DP), stream))
    return out


def beam_sums_device(p, gref, xiref, out, work, stream=None):
    """Beam::sums reduced on the device into `out` (17 doubles, device); work from beam_sums_workspace_bytes."""
    soa = beam_soa(p)
    check(lib().qz_beam_sums_device(p["x"].numel(), C.byref(soa), gref, xiref, _ptr(out), _ptr(work), stream))


def beam_sums_workspace_bytes(np_):
    return lib().qz_beam_sums_workspace_bytes(np_)
