"""CPU checks of the product boundary: the C-ABI library loads and exports every function the public
header declares (no compute call without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "quarz_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(qz_[a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("qz_sweep", "qz_sweep_batched", "qz_beam_deposit", "qz_beam_advance", "qz_tridiag_thomas_batched",
                 "qz_sim_create", "qz_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import quarz_amd
    lib = ctypes.CDLL(quarz_amd.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_last_error_is_empty_string_initially():
    import quarz_amd
    assert isinstance(quarz_amd.last_error(), str)


def test_set_block_validates_and_round_trips():
    """The launch-shape table is host state: no GPU call is made."""
    import quarz_amd
    for k in quarz_amd.KERNELS:
        before = quarz_amd.get_block(k)
        assert before == 256
        for bad in (0, 32, 100, 2048):
            try:
                quarz_amd.set_block(k, bad)
                raise AssertionError(f"{k}: {bad} accepted")
            except quarz_amd.QuarzError:
                pass
        quarz_amd.set_block(k, 128)
        assert quarz_amd.get_block(k) == 128
        quarz_amd.set_block(k, before)


def test_entry_points_reject_bad_arguments_before_any_gpu_call():
    """Argument checks of the kernel entry points return QZ_EINVAL with a message before any HIP call (the
    pointers below are never read)."""
    import ctypes as C
    import quarz_amd
    L = quarz_amd.lib()
    dummy = C.c_void_p(256)
    g = quarz_amd.RadialGridC()          # N = 0
    box = quarz_amd.BeamGridC(0.0, 0.01, 10, 0, 10, 0)
    soa = quarz_amd.BeamSoAC()
    q = (C.c_double * 2)(-1.0, 1.0)
    n = (C.c_int * 2)(5, 7)

    def dep(nbeams, nc):
        return L.qz_beam_deposit_beams(C.byref(g), C.byref(box), 0, nbeams, q, n, C.byref(soa), dummy, nc, dummy,
                                       None, 0, None, None)

    assert dep(0, 5) == -1 and "no beam" in quarz_amd.last_error()
    assert dep(2, 5) == -1 and "bad extent" in quarz_amd.last_error()       # N < 2
    g.N = 8
    assert dep(2, 7) == -1 and "nc must be" in quarz_amd.last_error()       # m = 0 needs 5 components
    n[1] = -3
    assert dep(2, 5) == -1 and "bad extent" in quarz_amd.last_error()       # negative particle count
    assert L.qz_beam_deposit(C.byref(g), C.byref(box), 0, -1.0, 5, C.byref(soa), dummy, 15, dummy, None) == -1
    assert L.qz_beam_advance(C.byref(g), C.byref(box), 0, -1.0, 0, 0, 5, C.byref(soa), dummy, 18, 1.0, 1.0, 1,
                             None) == -1
    assert "nc must be" in quarz_amd.last_error()
    assert L.qz_tridiag_thomas_batched(dummy, dummy, dummy, dummy, dummy, 0, 1, 1, 1, dummy, None) == -1
    assert "bad extent" in quarz_amd.last_error()
    assert L.qz_tridiag_thomas_batched(dummy, dummy, dummy, dummy, dummy, 4, 2, 4, 1, None, None) == -1
    assert "null argument" in quarz_amd.last_error()                         # no workspace
    assert L.qz_tridiag_thomas_batched(dummy, dummy, dummy, dummy, dummy, 4, 2, 2, 1, dummy, None) == -1
    assert "overlap" in quarz_amd.last_error()                               # system 1 starts inside system 0
    assert L.qz_tridiag_thomas_batched(dummy, dummy, dummy, dummy, dummy, 4, 3, 2, 4, dummy, None) == -1
    assert "overlap" in quarz_amd.last_error()                               # (2, 0) is (0, 1)
    assert L.qz_tridiag_workspace_doubles(5, 3) == 15


def test_sweep_rejects_unsupported_sizes_before_any_gpu_call():
    """qz_sweep checks every run before it touches the device: a zero-filled argument block has N = 0."""
    import ctypes as C
    import quarz_amd
    L = quarz_amd.lib()
    blob = (C.c_char * 4096)()           # larger than qz_sweep_args, all zero: modes 0, N 0
    assert L.qz_sweep(blob, None) == -2 and "radial cells" in quarz_amd.last_error()
    assert L.qz_sweep_batched(blob, 0, None) == -1 and "no runs" in quarz_amd.last_error()


def test_null_handles_are_errors():
    """Every qz_sim_* entry point answers a null run handle (or a null output) with QZ_EINVAL and a message,
    before any device work."""
    import ctypes as C
    import math
    import quarz_amd
    L = quarz_amd.lib()
    out = (C.c_double * 32)()
    for fn, args in [("qz_sim_step", ()), ("qz_sim_step_fields", ()), ("qz_sim_push_beams", ()),
                     ("qz_sim_run_pipelined", (2, 4, 2, None)), ("qz_sim_beam_sums", (0, out, None)),
                     ("qz_sim_beam_diag", (0, out)), ("qz_sim_write_fields", (b"/tmp/x.bin",)),
                     ("qz_sim_timing", (out,)), ("qz_sim_stats", (out,)), ("qz_sim_phase_ticks", (out,)),
                     ("qz_sim_snapshot", ()), ("qz_sim_rewind", ()), ("qz_sim_info", ((C.c_int * 10)(),))]:
        assert getattr(L, fn)(None, *args) == -1, fn
        assert "null" in quarz_amd.last_error(), (fn, quarz_amd.last_error())
    L.qz_sim_get.restype = C.c_long
    assert L.qz_sim_get(None, b"fld", out, 32) == -1 and "null" in quarz_amd.last_error()
    L.qz_sim_lost.restype = C.c_long
    assert L.qz_sim_lost(None) == -1 and "null" in quarz_amd.last_error()
    L.qz_sim_time.restype = C.c_double
    assert math.isnan(L.qz_sim_time(None)) and "null" in quarz_amd.last_error()


def test_noalloc_sweep_rejects_grids_beyond_its_scratch():
    """qz_sweep_noalloc keeps the run's workspace in a static array of the library sized for 2049 nodes: a larger
    grid is refused with QZ_ERANGE before any device call."""
    import ctypes as C
    import quarz_amd
    L = quarz_amd.lib()
    args = (C.c_char * 4096)()
    # qz_radial_grid leads the block: the r, inv_h and inv_V pointers, then N
    C.c_int.from_buffer(args, 3 * C.sizeof(C.c_void_p)).value = 100000
    assert L.qz_sweep_noalloc(args, None) == -2
    assert "scratch" in quarz_amd.last_error()


def test_message_entry_points_reject_bad_arguments():
    """qz_pack_message / qz_unpack_message check their arguments on the host before any device call."""
    import ctypes as C
    import quarz_amd
    L = quarz_amd.lib()
    assert L.qz_message_doubles(None, 1) == 0
    assert L.qz_pack_message(None, 1, 0, 0.0, 0, 1, None, None) == -1 and "null" in quarz_amd.last_error()
    assert L.qz_unpack_message(None, 1, None, None, None) == -1 and "null" in quarz_amd.last_error()
    args = (C.c_char * 4096)()   # a zeroed block: no workspace
    dummy = C.c_void_p(8)
    assert L.qz_pack_message(args, 1, 0, 0.0, 0, 0, dummy, None) == -1 and "ws" in quarz_amd.last_error()
    assert L.qz_unpack_message(args, 1, dummy, None, None) == -1 and "ws" in quarz_amd.last_error()
    assert 0 < C.sizeof(quarz_amd.SweepArgsHeadC) < L.qz_sweep_args_bytes() <= 4096


def test_batch_entry_points_reject_bad_arguments():
    """The batch's autotuner, worker cap, step and pipeline schedule answer null handles and bad counts with
    QZ_EINVAL and a message, before any device work."""
    import ctypes as C
    import quarz_amd
    L = quarz_amd.lib()
    cand = (C.c_int * 3)(1, 2, 4)
    groups = (C.c_int * 4)()
    est = (C.c_double * 2)()
    assert L.qz_batch_balance(None, cand, 3, 64, groups, est) == -1
    assert "bad arguments" in quarz_amd.last_error()
    assert L.qz_batch_set_workers(None, 4) == -1 and "null batch" in quarz_amd.last_error()
    assert L.qz_batch_step(None, None) == -1 and "null batch" in quarz_amd.last_error()
    assert L.qz_sim_step_batched(None, 0, None) == -1 and "no runs" in quarz_amd.last_error()
    assert L.qz_batch_create(None, 0, 1, None) == -1 and "no runs" in quarz_amd.last_error()
    nc = C.c_int(0)
    assert L.qz_sim_pipeline_timeline(None, None, 0, C.byref(nc)) == -1 and "null" in quarz_amd.last_error()


def test_batch_plan_on_known_estimates():
    """The batch autotuner's planning step on hand-made estimates (host only).  Four runs on 8 compute units:
    the longest (80 ms alone, 50 ms on 2 workgroups, 30 ms on 4) sets the one-workgroup-per-run makespan; the
    plan gives it 4 workgroups and list-schedules the other three on the 4 units left, for 30 ms; when extra
    workgroups do not shorten a run the plan keeps one per run; when no plan fits the units it is an error."""
    import numpy as np
    import quarz_amd
    est = np.array([[80.0, 50.0, 30.0], [30.0, 25.0, 20.0], [20.0, 15.0, 12.0], [10.0, 8.0, 7.0]])
    p = quarz_amd.batch_plan(est, (1, 2, 4), ncu=8)
    assert p["groups"] == [4, 1, 1, 1]
    assert p["est_ms"] == 30.0 and p["est_one_ms"] == 80.0
    # no speed-up from more workgroups: one each
    flat = np.array([[40.0, 45.0, 50.0]] * 3)
    p = quarz_amd.batch_plan(flat, (1, 2, 4), ncu=8)
    assert p["groups"] == [1, 1, 1] and p["est_ms"] == 40.0
    # two units for three runs: list scheduling, longest first
    p = quarz_amd.batch_plan(np.array([[30.0], [20.0], [15.0]]), (1,), ncu=2)
    assert p["groups"] == [1, 1, 1] and p["est_ms"] == 35.0
    # a run that cannot take one workgroup and no count that fits: an error
    try:
        quarz_amd.batch_plan(np.array([[-1.0, -1.0]]), (1, 2), ncu=8)
    except quarz_amd.QuarzError as e:
        assert "no plan" in str(e)
    else:
        raise AssertionError("an unplannable batch was accepted")


def test_laser_entry_points_reject_bad_arguments():
    """The laser envelope's entry points answer null handles, bad extents, schemes and carriers with an error code and
    a message, before any device work."""
    import ctypes as C
    import numpy as np
    import quarz_amd
    L = quarz_amd.lib()
    nodes = np.linspace(0.0, 1.0, 11)
    DP = C.POINTER(C.c_double)
    out = C.c_void_p()
    np_ = nodes.ctypes.data_as(DP)
    assert L.qz_laser_create(None, 11, 10, 0.0, 0.1, 1.0, 10.0, 1, 0, C.byref(out)) == -1
    assert "null" in quarz_amd.last_error()
    assert L.qz_laser_create(np_, 2, 10, 0.0, 0.1, 1.0, 10.0, 1, 0, C.byref(out)) == -1
    assert L.qz_laser_create(np_, 11, 10, 0.0, 0.1, 1.0, 10.0, 3, 0, C.byref(out)) == -1
    assert "scheme" in quarz_amd.last_error()
    assert L.qz_laser_create(np_, 11, 10, 0.0, 0.1, 1.0, 0.5, 1, 0, C.byref(out)) == -1
    assert "underdense" in quarz_amd.last_error()
    assert L.qz_laser_create(np_, 11, 10, 0.0, 0.0, 1.0, 10.0, 1, 0, C.byref(out)) == -1
    big = np.linspace(0.0, 1.0, 2100)
    assert L.qz_laser_create(big.ctypes.data_as(DP), 2100, 10, 0.0, 0.1, 1.0, 10.0, 1, 0, C.byref(out)) == -2
    assert L.qz_laser_step(None, None, None) == -1 and "null" in quarz_amd.last_error()
    assert L.qz_laser_gaussian(None, 0.1, 1.0, 1.0, 0.0, 0.0) == -1
    assert L.qz_laser_sums(None, None) == -1
    assert L.qz_laser_get(None, 0, None) == -1


def test_laser_deck_refused_by_the_step():
    """A deck with a laser (validation/laser_vacuum.in's keys) is refused by qz_sim_create before any device work:
    the step has no ponderomotive coupling, so it does not run such a deck without it."""
    import ctypes as C
    import quarz_amd
    L = quarz_amd.lib()
    deck = open(__file__.replace("test_capi.py", "decks/gain.in")).read() + "\nlaser.k0 = 20\nlaser.a0 = 0.1\n"
    h = C.c_void_p()
    assert L.qz_sim_create(deck.encode(), b"", C.byref(h)) == -1
    assert "ponderomotive" in quarz_amd.last_error()


def test_laser_batched_step_rejects_bad_arguments():
    """qz_laser_step_batched answers an empty batch and a null box with QZ_EINVAL before any device work."""
    import ctypes as C
    import quarz_amd
    L = quarz_amd.lib()
    assert L.qz_laser_step_batched(None, 0, None, None) == -1 and "no boxes" in quarz_amd.last_error()
    hs = (C.c_void_p * 2)(None, None)
    assert L.qz_laser_step_batched(hs, 2, None, None) == -1 and "null box" in quarz_amd.last_error()
