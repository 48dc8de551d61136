"""Benchmark of the quasi-static step on one MI355X: plasma particle-slices/s and beam pushes/s, FP64.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2] [--no-cpu]

A step is one quasi-static time step of the workload (quarz_amd/configs.py): beam deposit, the plasma sweep
over every slice (one persistent kernel), the beam push.  W untimed steps one at a time (at least one: the
first starts the leapfrog), then exactly K timed steps between device synchronisations.  The timed steps run
pipelined (Simulation.run_pipelined: up to --depth steps in flight, each split into --chunks slice ranges,
bit-identical to stepping one at a time); --depth 0 steps them one at a time.  --groups sets the workgroups of
each step's sweep (the defaults when pipelined: c2 6, so the 20 steps in flight hold 120 compute units; c3 16,
10 steps x 16; the eager steps run the same kernel).  Prints one JSON line:
  value          plasma particle-slices per second of the whole timed steps (wall clock)
  plain_value    the same run stepped one step at a time (--plain-steps steps), for comparison
  beam_pushes_per_s   live beam particles pushed per second of the whole timed steps
  roofline       the sweep kernel against MI355X FP64 vector peak: algorithmic flops, 679 per particle-slice +
                 508 per node-slice (m = 1; 361 / 84 for m = 0), kernel time from HIP events on the simulation's
                 stream; valu_frac: the FP64 VALU lane operations rocprof counts per launch
                 (profiles/sweep_fp64_valu.md) over the same time
  beam_roofline  the beam deposit and push against MI355X HBM bandwidth (8 TB/s), algorithmic bytes per step:
                 deposit 56 B read per live beam particle + bsrc (K M NC_B doubles) zeroed and written once;
                 push 56 B read + 48 B written per particle + fld (K M NC_F doubles) read once; times from HIP
                 events around the deposit (bsrc reset, live count, deposit) and the push launches
  cpu_baseline   the oracle's OpenMP build of the same step (the reference's loop structure) timed on this
                 host's cores on one full sweep of the same deck
The inputs (beam particles, grid, operators) are resident in HBM before the timed region; a cold plasma is
formed at the head of the box by the sweep kernel every step.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "plasma particle-slices/s & beam pushes/s (FP64, 1 GPU); batched vs plain"
FP64_PEAK = 78.6e12
HBM_PEAK = 8.0e12
FLOPS = {0: (361.0, 84.0), 1: (679.0, 508.0)}   # per particle-slice, per node-slice
# HBM bytes of one sweep launch (read + written), from rocprofv3 FETCH_SIZE and WRITE_SIZE passes over the
# same command, scaled by the counters' reading of a 1 GiB copy (profiles/sweep_c2_hbm_traffic.md,
# profiles/sweep_c3_counters.md)
# Keyed by (workload, workgroups per run); per whole sweep (every slice of one step).
MEASURED_TRAFFIC = {("c2", 1): 841.9e6, ("c2", 4): 20163.2e6, ("c2", 6): 21787.9e6, ("c3", 16): 699.5e9}
# FP64 VALU lane operations per sweep (every slice of one step), 64 x SQ_INSTS_VALU_FLOPS_FP64
# (profiles/sweep_fp64_valu.md): c2 one run, c4 the 33-run batched launch
MEASURED_VALU = {("c2", 1): 2.33e10, ("c2", 4): 2.481e10, ("c2", 6): 2.497e10, ("c3", 16): 5.066e11, ("c4", 1): 1.33e11}


def physical_cores():
    """Physical cores among the CPUs this process may run on (unique (package, core) pairs of /proc/cpuinfo)."""
    cpus = os.sched_getaffinity(0) if hasattr(os, "sched_getaffinity") else set(range(os.cpu_count() or 1))
    seen, cur = set(), {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if not line.strip():
                    if cur.get("processor") in cpus:
                        seen.add((cur.get("physical id"), cur.get("core id")))
                    cur = {}
                    continue
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                cur[k] = int(v) if k == "processor" else v
        if cur.get("processor") in cpus:
            seen.add((cur.get("physical id"), cur.get("core id")))
    except OSError:
        pass
    return max(1, len(seen)) if seen else len(cpus)


def _cpu_child(config, nthreads, nslices):
    """Runs in a child process (bench.py --cpu-child): one sweep of the oracle's OpenMP build on the deck,
    or its first nslices slices, with nthreads threads bound as the environment says.  Prints one JSON line."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    from quarz_amd import configs
    oracle_lib.set_threads(nthreads)
    s = oracle_lib.Sim(configs.deck(config), parallel=True)
    t0 = time.perf_counter()
    if nslices > 0:
        s.sweep_partial(nslices)
    else:
        s.step_fields()
    wall = time.perf_counter() - t0
    _, _, psl = s.timing()
    print(json.dumps({"wall": wall, "psl": psl, "K": s.K, "threads": nthreads}))


def cpu_baseline(config, nthreads, nslices=0):
    """The oracle's OpenMP build (the reference's loop structure: one parallel loop per kernel, atomic deposits,
    serial Thomas; g++ -O3 -ffp-contract=off) timed on this host in a child process, with OMP_PROC_BIND=close and
    OMP_PLACES=cores: nthreads threads and one thread.  MPI-style context: the one-thread rate times the cores
    at 90% efficiency (the reference's xi-pipelined MPI ranks scale near linearly, its OpenMP does not)."""
    import subprocess

    def run(n):
        env = dict(os.environ, OMP_NUM_THREADS=str(n), OMP_PROC_BIND="close", OMP_PLACES="cores")
        out = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child", config, str(n), str(nslices)],
                             env=env, capture_output=True, text=True, timeout=900, check=True)
        return json.loads(out.stdout.strip().splitlines()[-1])

    par = run(nthreads)
    one = run(1)
    rate, rate1 = par["psl"] / par["wall"], one["psl"] / one["wall"]
    what = f"the first {nslices} slices of one sweep" if nslices > 0 else f"one full sweep ({par['K']} slices)"
    return {"value": rate, "unit": "particle-slices/s", "cores": nthreads, "kind": "port",
            "sample": f"{what} of the same deck, {int(par['psl'])} particle-slices: {par['wall']:.2f} s on "
                      f"{nthreads} threads, {one['wall']:.2f} s on one (OpenMP atomics, OMP_PROC_BIND=close, "
                      f"OMP_PLACES=cores, g++ -O3 -ffp-contract=off)",
            "one_thread_value": rate1,
            "host_physical_cores": physical_cores(),
            "mpi_style_estimate": rate1 * nthreads * 0.90,
            "mpi_style_note": "context only: one-thread rate x cores x 0.90"}


def bench_batched(args):
    """c5: every step of all runs in one batched launch, against the same runs stepped one launch at a time."""
    import torch
    import quarz_amd
    from quarz_amd import configs
    warmup = max(args.warmup, 2)   # the leapfrog start-up step (eager), then the step that captures the graph
    runs = configs.c5_runs(args.runs, steps=warmup + args.steps + args.plain_steps + 1)
    sims = []
    t0 = time.perf_counter()
    for i, (deck, _) in enumerate(runs):
        sims.append(quarz_amd.Simulation(deck))
        if (i + 1) % 32 == 0:
            print(f"# set up {i + 1}/{len(runs)} runs ({time.perf_counter() - t0:.0f} s)", file=sys.stderr, flush=True)
    # launch shapes of the beam kernels, tuned on the run with the most beam particles
    big = max(range(len(runs)), key=lambda i: runs[i][1]["nb"])
    shapes = sims[big].autotune(cache=args.tune_cache)
    batch = quarz_amd.Batch(sims, graph=True)
    plan = None
    if args.balance:   # workgroups per run, from per-shape probes of the first slices
        tb = time.perf_counter()
        plan = batch.balance()
        plan["seconds"] = time.perf_counter() - tb
        print(f"# balanced: estimate {plan['est_ms']:.0f} ms (one workgroup per run {plan['est_one_ms']:.0f} ms) in "
              f"{plan['seconds']:.1f} s", file=sys.stderr, flush=True)
    te = time.perf_counter()
    kern_ms = batch.step()   # eager: the batched sweep launch time
    torch.cuda.synchronize()
    eager_wall = time.perf_counter() - te
    for _ in range(warmup - 1):
        batch.step()
    torch.cuda.synchronize()
    psl = pushes = 0.0
    t0 = time.perf_counter()
    graph_ms = 0.0
    for _ in range(args.steps):
        graph_ms += batch.step()
        for s in sims:
            st = s.stats()
            psl += st["particle_slices"]
            pushes += st["beam_pushed"]
        print(f"# batched step done ({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    replays = batch.replays
    batch.close()
    # the plain path: the same runs, one launch per run, one after the other
    ppsl = 0.0
    t1 = time.perf_counter()
    for _ in range(args.plain_steps):
        for i, s in enumerate(sims):
            s.step()
            ppsl += s.stats()["particle_slices"]
            if (i + 1) % 64 == 0:
                print(f"# plain: {i + 1}/{len(sims)} runs stepped ({time.perf_counter() - t1:.0f} s)", file=sys.stderr,
                      flush=True)
    torch.cuda.synchronize()
    pwall = time.perf_counter() - t1
    fp, fn = FLOPS[0]
    node_slices = sum(s.M * s.K for s in sims)
    flops = fp * psl / args.steps + fn * node_slices
    achieved = flops / (kern_ms / 1e3)
    Ns = sorted(s.M - 1 for s in sims)
    out = {
        "metric": METRIC,
        "value": psl / wall,
        "unit": "particle-slices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: beams sampled from per-run seeds, plasma loaded at the box head",
        "config": {"workload": f"c5: {configs.DESCRIPTION['c5']}", "runs": len(sims),
                   "cells_min_median_max": [Ns[0], Ns[len(Ns) // 2], Ns[-1]],
                   "slices_total": sum(s.K for s in sims),
                   "parallelism": ("workgroups per run from the batch autotuner: one-workgroup runs queued on the "
                                   "free compute units, multi-workgroup runs in a launch beside them") if plan else
                                  "one workgroup per run, one launch"},
        "beam_pushes_per_s": pushes / wall,
        "batched_sweep_ms": kern_ms,
        "graph_step_ms": graph_ms / args.steps,
        "graph_replays": replays,
        "eager_step_ms": eager_wall * 1e3,
        "launch_shapes": shapes,
        "balance": None if plan is None else {
            "groups_histogram": {str(g): plan["groups"].count(g) for g in sorted(set(plan["groups"]))},
            "multi_runs": [{"N": runs[i][1]["N"], "K": runs[i][1]["K"], "groups": g}
                           for i, g in enumerate(plan["groups"]) if g > 1],
            "est_ms": plan["est_ms"], "est_one_group_ms": plan["est_one_ms"], "tune_s": plan["seconds"]},
        "plain_value": ppsl / pwall if ppsl > 0 else None,
        "plain_steps": args.plain_steps,
        "batched_vs_plain": (psl / wall) / (ppsl / pwall) if ppsl > 0 else None,
        "roofline": {"bound": "fp64-valu", "achieved": achieved / 1e12, "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK, "traffic": None, "kernel": "sweep_kernel (batched)",
                     "flops_per_launch": flops,
                     "note": "FP64 vector work; MI355X's FP64 vector and matrix peaks are both 78.6 TFLOP/s"},
    }
    print(json.dumps(out))


def bench_gradient(args):
    """c4: the differentiable step.  Each timed step is one Adam iteration on the witness profile with the
    adjoint gradient: one step whose sweep is checkpointed every 64 slices, then the reverse xi sweep (each
    segment recomputed from its checkpoint) and the beam-side adjoints.  Beside it, at the initial profile: the
    central finite-difference gradient of the same step (2 S + 1 runs stepped in one batched launch) at eps 1e-6
    and 1e-7 and its agreement with the adjoint per component, and the same 2 S + 1 evaluations stepped one run
    at a time (batched vs plain)."""
    import numpy as np
    import torch
    from quarz_amd import configs, witness
    iters = args.steps
    S = configs.C4_SEGMENTS
    v0 = np.array(configs.c4_initial_profile())
    deck = configs.c4_deck(profile=v0)
    obj = witness.SpreadObjective(deck, beam=1, nseg=S, method="adjoint", every=64)
    shapes = obj.sims[0].autotune(cache=args.tune_cache)
    for _ in range(max(args.warmup, 1)):
        obj.value_and_grad(v0)
    torch.cuda.synchronize()
    # the check: central differences of the forward, batched
    fdo = witness.SpreadObjective(deck, beam=1, nseg=S, method="fd")
    J_adj, g_adj = obj.value_and_grad(v0)
    fd_check = {}
    fd_ms = None
    fd_errs = []
    for eps in (1e-6, 1e-7):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        J_fd, g_fd = fdo.value_and_grad(v0, eps=eps)
        torch.cuda.synchronize()
        if fd_ms is None:
            fd_ms = (time.perf_counter() - t0) * 1e3
        err = np.abs(g_fd - g_adj) / np.abs(g_adj).max()
        fd_errs.append(err)
        fd_check[str(eps)] = {"J": float(J_fd), "max_rel_err": float(err.max()), "median_rel_err": float(np.median(err)),
                              "components_within_1e-6": int((err < 1e-6).sum())}
    best = np.minimum(fd_errs[0], fd_errs[1])
    fd_batch_ms = fdo.last_batch_ms
    # the plain path: the 2 S + 1 evaluations of one finite-difference gradient, one run stepped at a time
    prof = [v0] + [v0 + d * e for e in np.eye(S) * 1e-6 for d in (1, -1)]
    ppsl = 0.0
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for s, p in zip(fdo.sims, prof):
        s.rewind()
        s.set_beam_profile(1, p)
        s.step()
        ppsl += s.stats()["particle_slices"]
    torch.cuda.synchronize()
    pwall = time.perf_counter() - t1
    fdo.close()
    # the timed Adam iterations on the adjoint gradient
    psl = 0.0
    fwd_ms = rev_ms = 0.0
    log = []

    def count(it, J):
        nonlocal psl, fwd_ms, rev_ms
        psl += obj.sims[0].stats()["particle_slices"]
        fwd_ms += obj.last_adjoint_ms[0]
        rev_ms += obj.last_adjoint_ms[1]
        log.append(J)
        if (it + 1) % 10 == 0:
            print(f"# Adam iteration {it + 1}/{iters}: J = {J:.6e}", file=sys.stderr, flush=True)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    v, hist = witness.optimise(obj, v0, iters=iters, lr=0.02, log=count)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    J_final = obj.value_and_grad(v)[0]
    sim = obj.sims[0]
    fp, fn = FLOPS[0]
    flops = fp * psl / iters + fn * sim.M * sim.K
    achieved = flops / (fwd_ms / iters / 1e3)
    out = {
        "metric": METRIC,
        "value": psl / wall,
        "unit": "particle-slices/s",
        "n_gpus": 1,
        "steps": iters,
        "warmup": max(args.warmup, 1),
        "ms_per_step": wall / iters * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: beams sampled from the deck's seeds, plasma loaded at the box head",
        "config": {"workload": f"c4: {configs.DESCRIPTION['c4']}", "M": sim.M, "slices": sim.K,
                   "checkpoint_every": 64, "parallelism": "one workgroup; forward, recomputation and reverse sweep"},
        "gradient": "hand-written adjoint (reverse xi sweep, checkpoints every 64 slices)",
        "adjoint_ms": {"checkpointed_forward": fwd_ms / iters, "reverse": rev_ms / iters, "per_gradient": wall / iters * 1e3,
                       "reverse_kernel_phases_last": {k: v / 1e5 for k, v in obj.sims[0].adjoint_ticks().items()}},
        "fd_check": {"by_eps": fd_check, "max_rel_err_best_of_two_eps": float(best.max()),
                     "batched_fd_gradient_ms": fd_ms, "batched_sweep_ms": fd_batch_ms,
                     "note": "J jumps where a plasma particle crosses a node (the dxi-derivative deposit uses the shape's "
                             "derivative), so a difference straddling one is off; the components agree where neither "
                             "point straddles a jump"},
        "adam": {"iterations": iters, "lr": 0.02, "J_initial": hist[0], "J_final": float(J_final),
                 "profile_final": [float(x) for x in v]},
        "launch_shapes": shapes,
        "fd_plain_particle_slices_per_s": ppsl / pwall,
        "fd_gradient_batched_vs_plain": pwall / (fd_ms / 1e3),
        "adjoint_vs_batched_fd": (fd_ms / 1e3) / (wall / iters),
        "roofline": {"bound": "fp64-valu", "achieved": achieved / 1e12, "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK, "traffic": None, "kernel": "sweep_kernel (checkpointed forward)",
                     "flops_per_launch": flops,
                     "note": "FP64 vector work; MI355X's FP64 vector and matrix peaks are both 78.6 TFLOP/s"},
    }
    obj.close()
    print(json.dumps(out))


def calibrate_copy():
    """bench.py --calibrate-copy: one plain device copy of known size (2**27 float64: 1 GiB read, 1 GiB
    written), the reference reading for the HBM byte counters (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE) that
    scale the kernels' counter bytes to HBM bytes."""
    import torch
    n = 1 << 27
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibration_copy_bytes": n * 8}))


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--cpu-child":
        _cpu_child(sys.argv[2], int(sys.argv[3]), int(sys.argv[4]))
        return
    if len(sys.argv) > 1 and sys.argv[1] == "--calibrate-copy":
        calibrate_copy()
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (c2: the workload's 20 time steps; c4: 50 Adam iterations; c5: 3, after the "
                         "2 warm-up steps: the workload's 5)")
    ap.add_argument("--warmup", type=int, default=2)   # the second step is a regular one: its kernel times are reported
    ap.add_argument("--config", default="c2")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--runs", type=int, default=256, help="c5: number of independent runs")
    ap.add_argument("--plain-steps", type=int, default=1, help="c5: steps of the one-launch-per-run comparison")
    ap.add_argument("--no-balance", dest="balance", action="store_false",
                    help="c5: one workgroup per run (skip the batch autotuner)")
    ap.add_argument("--depth", type=int, default=-1,
                    help="time steps in flight (-1: min(steps, 32); 0: one step at a time, no pipeline)")
    ap.add_argument("--chunks", type=int, default=96, help="slice ranges per step of the pipelined run")
    ap.add_argument("--groups", type=int, default=-1,
                    help="workgroups per run of the sweep (0: the run's own choice; -1: 6 for a pipelined c2 run, "
                         "whose 20 steps in flight then hold 120 of the 256 compute units, 16 for a pipelined c3 "
                         "run (10 x 16), else the run's own)")
    ap.add_argument("--tune-cache", default=None,
                    help="JSON file of launch shapes: reused for the same device and shapes instead of timing")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = {"c1": 10, "c3": 10, "c4": 50, "c5": 3}.get(args.config, 20)
    if args.config == "c1" and args.chunks == 96:
        args.chunks = 48   # ranges of 33 slices: the measured best for the short box (32: 65.8e6, 48: 70.3e6, 64: 65.8e6)
    if args.config == "c3":
        args.warmup = min(args.warmup, 1)
        if args.groups < 0 and args.depth != 0:   # the measured best shape of the pipelined long box (96 ranges)
            args.groups = 16   # 10 steps x 16 workgroups in flight: 160 of the 256 compute units
    # The pipelined run wants one hardware queue per step in flight; HIP's default is 4 per process, read when
    # HIP starts (so set before torch is imported), never above 32.  The batched launches (c4, c5) keep the
    # default: with 32 queues their long persistent launch ran about 1.5x slower (c4: 412 against 276 ms).
    if args.config not in ("c4", "c5") and args.depth != 0:
        want = min(32, max(4, (min(args.steps, 32) if args.depth < 0 else args.depth) + 4))
        if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < want:
            os.environ["GPU_MAX_HW_QUEUES"] = str(want)
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported (independent runs would "
              f"shard across GPUs by run, one large run by slice blocks)", file=sys.stderr)
        sys.exit(2)
    if args.config == "c5":
        bench_batched(args)
        return
    if args.config == "c4":
        bench_gradient(args)
        return
    import torch
    import quarz_amd
    from quarz_amd import configs
    deck = configs.deck(args.config)
    pipelined = args.depth != 0
    warmup = max(args.warmup, 1) if pipelined else args.warmup   # the pipeline continues a started run
    depth = min(args.steps, 32) if args.depth < 0 else args.depth
    total = warmup + args.steps + 1
    t_setup = time.perf_counter()
    if args.groups < 0:
        args.groups = 6 if (pipelined and args.config == "c2") else 0
    over = {"sweep.groups": args.groups} if args.groups > 0 else {}
    sim = quarz_amd.Simulation(deck, {"time.steps": total, **over})
    print(f"# {args.config} set up in {time.perf_counter() - t_setup:.1f} s ({sim.groups} workgroups per run)",
          file=sys.stderr, flush=True)
    modes = sim.m1
    group_times = None
    if sim.groups > 1 and args.groups == 0:   # workgroups per run, timed on the first slices of the first step
        group_times = sim.tune_groups([8, 16, 32], nslices=64)
    shapes = sim.autotune(cache=args.tune_cache)   # launch shapes of the beam kernels for this run
    # warm-up steps, one at a time (the first starts the leapfrog); the last one's HIP-event kernel times
    # give the per-launch figures (sweep, beam deposit, beam push)
    eager = None
    for i in range(warmup):
        sim.step()
        eager = sim.stats()
        print(f"# warm-up step {i + 1}/{warmup}: sweep {eager['sweep_ms']:.0f} ms", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    psl = pushes = dep_particles = 0.0
    sweep_ms = beam_ms = dep_ms = 0.0
    t0 = time.perf_counter()
    timed_sweep_ms, launches = None, 1
    if pipelined:
        psl, pushes = sim.run_pipelined(args.steps, chunks=args.chunks, depth=depth)
        st = sim.stats()
        dep_particles = st["beam_deposited"] * args.steps
        # kernel time of the timed steps: every range launch bracketed by HIP events on its own stream
        timed_sweep_ms, launches = st["sweep_ms"], int(st["sweep_launches"])
    else:
        for i in range(args.steps):
            sim.step()
            st = sim.stats()
            print(f"# step {i + 1}/{args.steps}: sweep {st['sweep_ms']:.0f} ms ({time.perf_counter() - t0:.1f} s)",
                  file=sys.stderr, flush=True)
            psl += st["particle_slices"]
            pushes += st["beam_pushed"]
            dep_particles += st["beam_deposited"]
            sweep_ms += st["sweep_ms"]
            beam_ms += st["beam_push_ms"]
            dep_ms += st["beam_dep_ms"]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    if pipelined:   # per-launch kernel times from the last eager step
        sweep_ms = eager["sweep_ms"] * args.steps
        beam_ms = eager["beam_push_ms"] * args.steps
        dep_ms = eager["beam_dep_ms"] * args.steps
    # the same run stepped one step at a time, for comparison
    plain_value = None
    if pipelined and args.plain_steps > 0:
        ps = quarz_amd.Simulation(deck, {"time.steps": args.plain_steps + 2, **over})
        ps.step()
        torch.cuda.synchronize()
        tp = time.perf_counter()
        ppsl = 0.0
        for _ in range(args.plain_steps):
            ps.step()
            ppsl += ps.stats()["particle_slices"]
        torch.cuda.synchronize()
        plain_value = ppsl / (time.perf_counter() - tp)
    fp, fn = FLOPS[modes]
    flops_per_sweep = fp * psl / args.steps + fn * sim.M * sim.K
    # the dominant kernel's rate: algorithmic flops per launch over the average launch time of the timed steps
    # (pipelined: flops of one step's slice ranges over the sum of their launch times)
    t_kernel = (timed_sweep_ms if timed_sweep_ms is not None else sweep_ms / args.steps) / 1e3
    achieved = flops_per_sweep / t_kernel
    mkey = (args.config, sim.groups)   # the counter readings of this kernel and shape
    out = {
        "metric": METRIC,
        "value": psl / wall,
        "unit": "particle-slices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: beams sampled from the deck's seeds, plasma loaded at the box head",
        "config": {"workload": f"{args.config}: {configs.DESCRIPTION[args.config]}", "M": sim.M, "slices": sim.K,
                   "plasma_particles_per_slice": int(round(psl / args.steps / max(sim.K, 1))),
                   "beam_particles": int(round(pushes / args.steps)),
                   "parallelism": (f"one run per GPU; {depth} time steps in flight, {args.chunks} slice ranges per step, "
                                   f"{sim.groups} workgroup(s) per step's sweep"
                                   if pipelined else f"one run per GPU, one step at a time, {sim.groups} workgroup(s)")},
        "plain_value": plain_value,
        "beam_pushes_per_s": pushes / wall,
        "sweep_ms": sweep_ms / args.steps,
        "timed_sweep_launch_ms_per_step": timed_sweep_ms,
        "beam_deposit_ms": dep_ms / args.steps,
        "beam_push_ms": beam_ms / args.steps,
        "sweep_particle_slices_per_s": psl / (sweep_ms / 1e3),
        "launch_shapes": shapes,
        "sweep_groups": sim.groups,
        "sweep_group_times_ms": group_times,
        "roofline": {"bound": "fp64-valu", "achieved": achieved / 1e12, "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK,
                     "traffic": (MEASURED_TRAFFIC[mkey] / launches if mkey in MEASURED_TRAFFIC else None),
                     "traffic_per_sweep": MEASURED_TRAFFIC.get(mkey),
                     "kernel": "sweep_kernel" if sim.groups == 1 else f"sweep_multi_kernel ({sim.groups} workgroups)",
                     "launches_per_step": launches,
                     "avg_launch_ms": t_kernel * 1e3 / launches,
                     "flops_per_launch": flops_per_sweep / launches,
                     "valu_fp64_lane_ops_per_sweep": MEASURED_VALU.get(mkey),
                     "valu_frac": (MEASURED_VALU[mkey] / t_kernel / FP64_PEAK
                                   if mkey in MEASURED_VALU and t_kernel > 0 else None),
                     "aggregate_achieved": flops_per_sweep * args.steps / wall / 1e12,
                     "aggregate_frac": flops_per_sweep * args.steps / wall / FP64_PEAK,
                     "note": "FP64 vector work, no dense contraction (no MFMA); MI355X's FP64 vector peak is 78.6 "
                             "TFLOP/s. achieved: algorithmic flops per sweep launch over the average launch time "
                             "of the timed steps (HIP events around every launch on its stream); aggregate_achieved: "
                             "every sweep of the timed steps over their wall time (pipelined steps overlap on "
                             "several CUs)"},
    }
    nb_dep, nb_push = dep_particles / args.steps, pushes / args.steps
    ncf, ncb = (18, 15) if modes else (6, 5)
    kmn = 8.0 * sim.K * sim.M
    parts = {"deposit": (56.0 * nb_dep + 2.0 * kmn * ncb, dep_ms), "push": (104.0 * nb_push + kmn * ncf, beam_ms)}
    beam = {"bound": "hbm", "peak": HBM_PEAK / 1e9, "unit": "GB/s"}
    tot_b = tot_t = 0.0
    for name, (b, ms) in parts.items():
        t = ms / args.steps / 1e3
        tot_b += b
        tot_t += t
        beam[name] = {"achieved": b / t / 1e9 if t > 0 else None, "frac": b / t / HBM_PEAK if t > 0 else None,
                      "bytes_per_step": b}
    beam["achieved"] = tot_b / tot_t / 1e9 if tot_t > 0 else None
    beam["frac"] = tot_b / tot_t / HBM_PEAK if tot_t > 0 else None
    out["beam_roofline"] = beam
    if not args.no_cpu:
        # the GPU box's share of the host is 16 cores per GPU
        # a bounded sample for the long box: the first 200 of its 16001 slices
        out["cpu_baseline"] = cpu_baseline(args.config, min(16, physical_cores()), 200 if args.config == "c3" else 0)
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
