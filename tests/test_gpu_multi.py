"""GPU parity of the sweep run by several cooperating workgroups (qz_sweep_args.groups > 1): bit-identical to the
serial oracle, on small decks with a forced workgroup count (uneven particle shares, one and two mobile species,
m = 0 and m = 1) and at the long-box workload's shape (1024 stretched cells, electrons and mobile ions, 32768
plasma particles per slice, 8 workgroups) in a short box."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qa():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import quarz_amd
    quarz_amd.lib()
    return quarz_amd


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a, b)


def _report(name, a, b):
    a, b = np.asarray(a), np.asarray(b)
    bad = np.nonzero(a != b)
    n = len(bad[0])
    first = tuple(int(x[0]) for x in bad) if n else None
    return f"{name}: {n} of {a.size} differ; first at {first}: {a[first] if n else ''} vs {b[first] if n else ''}"


def _compare(qa, g, o, steps):
    for _ in range(steps):
        g.step()
        o.step()
    for w in ("fld", "diag", "bsrc"):
        a, b = getattr(g, w), getattr(o, w)
        assert _same(a, b), _report(w, a, b)
    for bi in range(o.nbeams):
        for k in qa.BEAM_ARRAYS:
            a, b = g.get(f"beam{bi}.{k}"), o.get(f"beam{bi}.{k}")
            assert _same(a, b), _report(f"beam{bi} {k}", a, b)
    for sp in range(o.nspecies):
        for k in ("x", "y", "px", "py", "dl", "w"):
            a, b = g.get(f"sp{sp}.{k}"), o.get(f"sp{sp}.{k}")
            assert _same(a, b), _report(f"species {sp} {k}", a, b)
    assert g.lost() == o.lost()


M1_DECK = {"modes": 1, "xi.min": -5, "grid.dr": 0.04, "driver.n0": 0.5, "driver.x0": 0.05, "driver.nparticles": 20000,
           "driver.sigma_r": 0.5, "driver.rigid": 0, "electrons.ppc": 1, "time.steps": 2, "time.dt": 5}
ION_DECK = {"xi.min": -7.3, "time.steps": 2, "driver.nparticles": 40000, "witness.nparticles": 20000}


@pytest.mark.parametrize("case,groups", [("m1", 2), ("m1", 3), ("ions", 3), ("ions", 4)])
def test_forced_groups_equal_oracle(qa, oracle, decks, case, groups):
    """m = 1 with 1600 mobile electrons on 2 or 3 workgroups (shares of 1024, 576 / 1024, 576, 0), and the m = 0
    ion-motion shape (3936 electrons and ions) on 3 or 4 workgroups: three steps each, bit for bit."""
    deck, ov = (f"{decks}/linear.in", M1_DECK) if case == "m1" else (f"{decks}/pwfa_ion_motion.in", ION_DECK)
    g = qa.Simulation(deck, dict(ov, **{"sweep.groups": groups}))
    o = oracle.Sim(deck, ov)
    _compare(qa, g, o, 3)


def test_c3_shape_bit_exact(qa, oracle, tmp_path):
    """The long-box workload's plasma and grid (m = 0 + 1, 1024 stretched cells to R = 12, electrons and mobile
    ions with 16384 particles per slice each, Picard 3) on 8 workgroups, in a box of 301 slices with a short
    offset driver and a witness: three steps, bit for bit."""
    from quarz_amd import configs
    deck = configs.c3_deck(workdir=str(tmp_path), xi_min=-1.5, steps=2, nd=100000, nw=20000, driver_xi0=-0.6,
                           witness_xi0=-1.2)
    ov = {"driver.sigma_xi": 0.2, "driver.x0": 0.05, "driver.n0": 2, "witness.sigma_xi": 0.05}
    g = qa.Simulation(deck, ov)
    o = oracle.Sim(deck, ov)
    assert g.M == 1025 and g.K == 301
    _compare(qa, g, o, 3)


def test_batched_multi_runs_equal_plain(qa, decks):
    """Two runs of several workgroups each in one batched launch: each equals its own plain launch."""
    ovs = [dict(M1_DECK, **{"sweep.groups": 2}), dict(M1_DECK, **{"sweep.groups": 3, "xi.min": -4, "grid.dr": 0.05})]
    plain = [qa.Simulation(f"{decks}/linear.in", ov) for ov in ovs]
    batch = [qa.Simulation(f"{decks}/linear.in", ov) for ov in ovs]
    for _ in range(3):
        for s in plain:
            s.step()
        qa.step_batched(batch)
    for i, (a, b) in enumerate(zip(batch, plain)):
        for w in ("fld", "diag", "bsrc"):
            assert _same(getattr(a, w), getattr(b, w)), (i, w)


@pytest.mark.parametrize("deck,ov,groups,steps,chunks,depth", [
    # a slow witness (gamma 3) slipping across several ranges per step, m = 0, 2 workgroups
    ("gain.in", {"time.steps": 4, "driver.nparticles": 20000, "xi.min": -8, "witness.gamma": 3,
                 "witness.sigma_xi": 1.0, "witness.xi0": -4, "witness.n0": 0.001}, 2, 3, 32, 3),
    # m = 1 mobile plasma, 3 workgroups, four steps in flight
    ("betatron.in", {"modes": 1, "witness.x0": 0.05, "witness.nsym": 4, "time.steps": 6}, 3, 5, 6, 4),
])
def test_pipelined_groups_equal_plain(qa, decks, deck, ov, groups, steps, chunks, depth):
    """Steps in flight whose sweeps run on several workgroups each (ordinary launches on the steps' streams,
    every workgroup resident) against the same steps on one workgroup, one at a time: fields, diagnostics, beam
    sources and every beam particle bit for bit."""
    a = qa.Simulation(f"{decks}/{deck}", ov)
    b = qa.Simulation(f"{decks}/{deck}", {**ov, "sweep.groups": groups})
    assert a.groups == 1 and b.groups == groups
    a.step()
    b.step()
    for _ in range(steps):
        a.step()
    b.run_pipelined(steps, chunks=chunks, depth=depth)
    for w in ("fld", "diag", "bsrc"):
        ga, gb = getattr(b, w), getattr(a, w)
        assert _same(ga, gb), _report(f"pipelined {w}", ga, gb)
    for bi in range(a.nbeams):
        for k in qa.BEAM_ARRAYS:
            ga, gb = b.get(f"beam{bi}.{k}"), a.get(f"beam{bi}.{k}")
            assert _same(ga, gb), _report(f"pipelined beam{bi}.{k}", ga, gb)
    assert b.lost() == a.lost()


# ------------------------------------------------------------------------------------------ mixed batches
GAIN_RUNS = [{"xi.min": -5, "witness.xi0": -4},
             {"xi.min": -6, "witness.xi0": -4.5, "grid.type": "uniform", "grid.dr": 0.02, "electrons.ppc": 2},
             {"xi.min": -4.5, "witness.xi0": -4, "grid.type": "uniform", "grid.dr": 0.08, "electrons.ppc": 8},
             {"xi.min": -5.5, "witness.xi0": -4.5, "electrons.ppc": 2},
             {"xi.min": -4, "witness.xi0": -3.5, "grid.type": "uniform", "grid.dr": 0.04}]
GAIN_BASE = {"time.steps": 5, "driver.nparticles": 20000, "witness.nparticles": 5000}


def _batch_vs_plain(qa, deck, batch_ovs, plain_ovs, steps, workers=0, balance=None):
    """A graph Batch against the same decks stepped one run at a time (on one workgroup where the run fits one),
    bit for bit after every step (the eager start-up step, the replays, the first of which captures, and the
    eager last step)."""
    plain = [qa.Simulation(deck, ov) for ov in plain_ovs]
    runs = [qa.Simulation(deck, ov) for ov in batch_ovs]
    batch = qa.Batch(runs, graph=True, workers=workers)
    try:
        plan = batch.balance(**balance) if balance else None
        for st in range(steps):
            for s in plain:
                s.step()
            batch.step()
            for i, (a, b) in enumerate(zip(runs, plain)):
                for w in ("fld", "diag", "bsrc"):
                    ga, gb = getattr(a, w), getattr(b, w)
                    assert _same(ga, gb), _report(f"step {st} run {i} (groups {a.groups}) {w}", ga, gb)
                for bi in range(a.nbeams):
                    for k in qa.BEAM_ARRAYS:
                        ga, gb = a.get(f"beam{bi}.{k}"), b.get(f"beam{bi}.{k}")
                        assert _same(ga, gb), _report(f"step {st} run {i} beam{bi}.{k}", ga, gb)
                assert a.lost() == b.lost()
                assert a.t == b.t
        replays = batch.replays
    finally:
        batch.close()
    return runs, plan, replays


def test_mixed_batch_queue_equals_plain(qa, decks):
    """Five runs of different grids and slice counts in one graph Batch: two on several workgroups (2 and 3, one
    launch beside the other) and three on one workgroup each, taken from the queue by two workers.  Every run
    equals its own one-workgroup plain run after every step."""
    ovs = [dict(GAIN_BASE, **o) for o in GAIN_RUNS]
    forced = [dict(o) for o in ovs]
    forced[0]["sweep.groups"] = 2
    forced[3]["sweep.groups"] = 3
    runs, _, replays = _batch_vs_plain(qa, f"{decks}/gain.in", forced, ovs, 6, workers=2)
    assert [r.groups for r in runs] == [2, 1, 1, 3, 1]
    assert replays == 4


def test_one_group_batch_queue_equals_plain(qa, decks):
    """One-workgroup runs only, one worker: the queue runs them one after another, longest first."""
    ovs = [dict(GAIN_BASE, **o) for o in GAIN_RUNS[:3]]
    _, _, replays = _batch_vs_plain(qa, f"{decks}/gain.in", ovs, ovs, 6, workers=1)
    assert replays == 4


def test_balanced_batch_equals_plain(qa, decks):
    """The batch autotuner: every shape probed with one to four workgroups, the plan applied to the runs, then
    the steps; the plan is no worse than one workgroup per run by its own estimate, and every run equals its
    plain run."""
    ovs = [dict(GAIN_BASE, **o) for o in GAIN_RUNS]
    runs, plan, replays = _batch_vs_plain(qa, f"{decks}/gain.in", ovs, ovs, 6,
                                          balance={"candidates": (1, 2, 3, 4), "nslices": 64})
    print("plan:", plan)
    assert plan["groups"] == [r.groups for r in runs]
    assert all(g in (1, 2, 3, 4) for g in plan["groups"])
    assert 0 < plan["est_ms"] <= plan["est_one_ms"] * (1 + 1e-12)
    assert replays == 4


def test_mixed_batch_two_launches_equals_plain(qa, decks):
    """A multi-workgroup run with more than 1024 plasma particles per workgroup (8 per cell on the gain grid, over
    4096, on 2 workgroups: 4 particles per thread) cannot share the one mixed launch, so its sweep goes to a
    launch of its own beside the queue's: every run still equals its plain run after every step."""
    ovs = [dict(GAIN_BASE, **o) for o in GAIN_RUNS[:3]]
    ovs[0]["electrons.ppc"] = 8
    forced = [dict(o) for o in ovs]
    forced[0]["sweep.groups"] = 2
    runs, _, replays = _batch_vs_plain(qa, f"{decks}/gain.in", forced, ovs, 6, workers=1)
    assert [r.groups for r in runs] == [2, 1, 1]
    assert replays == 4


def test_c3_shape_pipelined_equals_plain(qa, tmp_path):
    """The long-box workload's plasma and grid (1024 stretched cells, electrons and mobile ions, 32768 particles
    per slice, 8 workgroups) in a short box: four steps with three in flight over six ranges each equal the same
    steps one at a time, bit for bit (fields, diagnostics, beam sources, every beam particle)."""
    from quarz_amd import configs
    deck = configs.c3_deck(workdir=str(tmp_path), xi_min=-1.5, steps=5, nd=100000, nw=20000, driver_xi0=-0.6,
                           witness_xi0=-1.2)
    ov = {"driver.sigma_xi": 0.2, "driver.x0": 0.05, "driver.n0": 2, "witness.sigma_xi": 0.05}
    a = qa.Simulation(deck, ov)
    b = qa.Simulation(deck, ov)
    assert a.groups == 8 and b.groups == 8
    a.step()
    b.step()
    for _ in range(4):
        a.step()
    b.run_pipelined(4, chunks=6, depth=3)
    for w in ("fld", "diag", "bsrc"):
        ga, gb = getattr(b, w), getattr(a, w)
        assert _same(ga, gb), _report(f"pipelined {w}", ga, gb)
    for bi in range(a.nbeams):
        for k in qa.BEAM_ARRAYS:
            ga, gb = b.get(f"beam{bi}.{k}"), a.get(f"beam{bi}.{k}")
            assert _same(ga, gb), _report(f"pipelined beam{bi}.{k}", ga, gb)
    assert b.lost() == a.lost()
