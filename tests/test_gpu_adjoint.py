"""The device adjoint of one m = 0 step (qz_sweep_adjoint and the beam-side adjoints, driven by
qz_sim_spread_grad) against the oracle's adjoint (adjoint_oracle.inc, itself pinned by central differences in
test_adjoint_oracle.py) and against central differences of the device forward.

The device forward is bit-identical to the oracle's, so both adjoints differentiate the same trajectory; they
differ only by the rounding of their own sums (the device sums node contributions with atomics), so they agree
to a relative 1e-9 of the gradient's largest component.  The checkpoint interval changes nothing but that
rounding."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NSEG = 16


@pytest.fixture(scope="module")
def qa():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import quarz_amd
    quarz_amd.lib()
    return quarz_amd


@pytest.fixture(scope="module")
def oracle():
    import oracle_lib
    return oracle_lib


@pytest.fixture(scope="module")
def c4_small(tmp_path_factory):
    from quarz_amd import configs
    return configs.c4_deck(nd=5000, nw=4000, workdir=str(tmp_path_factory.mktemp("c4")))


@pytest.mark.parametrize("regime", ["weak", "blowout"])
def test_adjoint_equals_oracle_adjoint(qa, oracle, c4_small, regime):
    """Weak driver (n_b 0.05) and the workload's driver (n_b 2, electrons crossing the axis), box [-10.5, 0]."""
    ov = {"xi.min": -10.5}
    if regime == "weak":
        ov["driver.n0"] = 0.05
    o = oracle.Sim(c4_small, ov)
    Jo, go = o.spread_grad(1, NSEG)
    s = qa.Simulation(c4_small, ov)
    assert s.profile_size(1) == NSEG and s.profile_size(0) == 0
    J, g, ms = s.spread_grad(1, every=64)
    print(f"{regime}: J {J:.15e} forward {ms[0]:.1f} ms reverse {ms[1]:.1f} ms, reverse-sweep phases (ms):",
          {k: round(v / 1e5, 2) for k, v in s.adjoint_ticks().items()})
    assert J == pytest.approx(Jo, rel=1e-13)
    assert np.abs(g - go).max() <= 1e-9 * np.abs(go).max(), (g, go)
    # the step itself is the forward's: fields and the pushed beams equal the oracle's after its step
    assert np.array_equal(s.fld, o.fld)
    for k in qa.BEAM_ARRAYS:
        assert np.array_equal(s.get(f"beam1.{k}"), o.get(f"beam1.{k}")), k


def test_adjoint_checkpoint_interval(qa, c4_small):
    """Segments of 7 slices, 64 slices and the whole box give the same gradient (to rounding)."""
    ov = {"xi.min": -10.5}
    grads = []
    for every in (7, 64, 100000):
        s = qa.Simulation(c4_small, ov)
        J, g, _ = s.spread_grad(1, every=every)
        grads.append(g)
    for g in grads[1:]:
        assert np.abs(g - grads[0]).max() <= 1e-10 * np.abs(grads[0]).max()


def test_adjoint_equals_device_finite_differences(qa, c4_small):
    """Central differences of the device forward (2 S + 1 runs in one batched step).  J jumps where a plasma
    particle crosses a node (the dxi-derivative deposit's shape changes there), so each component is compared
    at eps = 1e-6 and 1e-7 and must agree at one of them; most agree at 1e-6."""
    from quarz_amd import configs
    from quarz_amd.witness import SpreadObjective
    ov = {"xi.min": -10.5, "driver.n0": 0.05}
    s = qa.Simulation(c4_small, ov)
    J, g, _ = s.spread_grad(1)
    prof = np.array(configs.c4_initial_profile())
    obj = SpreadObjective(c4_small, beam=1, nseg=NSEG, overrides=ov, method="fd")
    try:
        errs = []
        for eps in (1e-6, 1e-7):
            J0, fd = obj.value_and_grad(prof, eps=eps, method="fd")
            assert J0 == pytest.approx(J, rel=1e-13)
            errs.append(np.abs(fd - g) / np.abs(g).max())
        best = np.minimum(errs[0], errs[1])
        assert best.max() < 5e-6, errs
        assert (errs[0] < 1e-6).sum() >= 12, errs[0]
    finally:
        obj.close()
