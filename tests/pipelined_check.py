"""Child process of test_pipelined_with_many_hardware_queues: pipelined steps against plain steps, with the
hardware-queue count the bench uses (set by the parent before HIP starts here).  Exit status 0 when every
array is bit-identical."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import quarz_amd  # noqa: E402
from quarz_amd import configs  # noqa: E402

DECKS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "decks")
CASES = [
    (os.path.join(DECKS, "gain.in"), {"time.steps": 4, "driver.nparticles": 20000, "xi.min": -8, "witness.gamma": 3,
                                      "witness.sigma_xi": 1.0, "witness.xi0": -4, "witness.n0": 0.001}, 3, 32, 3),
    # the bench's c2 setting: two warm-up steps, then the workload's 20 time steps with 20 in flight over 96
    # slice ranges, each step's sweep on 4 workgroups (the run has 23 steps, so the last timed step is followed by
    # a push like every other); the plain steps run on one workgroup
    (configs.deck("c2"), {"time.steps": 23}, 20, 96, 20, {"sweep.groups": 4}),
]
bad = 0
for deck, ov, steps, chunks, depth, *rest in CASES:
    a = quarz_amd.Simulation(deck, ov)
    b = quarz_amd.Simulation(deck, {**ov, **(rest[0] if rest else {})})
    warm = 2 if depth == 20 else 1
    for _ in range(warm):
        a.step()
        b.step()
    for _ in range(steps):
        a.step()
    b.run_pipelined(steps, chunks=chunks, depth=depth)
    same = all(np.array_equal(getattr(a, w), getattr(b, w)) for w in ("fld", "diag", "bsrc"))
    for bi in range(a.nbeams):
        for k in quarz_amd.BEAM_ARRAYS:
            same = same and np.array_equal(a.get(f"beam{bi}.{k}"), b.get(f"beam{bi}.{k}"))
    print(f"{ov} {rest} steps {steps} chunks {chunks} depth {depth}: {'equal' if same else 'DIFFERENT'}", flush=True)
    bad += not same
sys.exit(1 if bad else 0)
