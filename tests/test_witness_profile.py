"""The differentiable workload: a flat-top witness whose weights follow a piecewise-linear profile.

CPU tests pin the oracle's profile weights against closed forms; GPU tests check the profile, snapshot and
rewind, the batched objective and its finite-difference gradient against the oracle and against plain runs.
"""
import numpy as np
import pytest

from quarz_amd import configs

S = configs.C4_SEGMENTS
SMALL = {"nd": 20000, "nw": 10000}


def _deck(profile, **kw):
    return configs.c4_deck(profile=profile, **{**SMALL, **kw})


def _short(deck):
    """1200 slices instead of 2000: the driver at -2 and the witness over [-10, -6] stay in the box."""
    assert "xi.min = -19.99\n" in deck
    return deck.replace("xi.min = -19.99\n", "xi.min = -11.99\n")


def _no_profile(deck):
    return "\n".join(l for l in deck.splitlines() if not l.startswith("witness.pwl_weights"))


def test_c4_deck_shape(oracle):
    o = oracle.Sim(configs.c4_deck(nd=1000, nw=1000))
    assert (o.M, o.K, o.m1, o.nbeams) == (513, 2000, 0, 2)


def test_profile_of_ones_is_the_flat_top(oracle):
    """Values 1 everywhere give the flat-top beam bit for bit (weight = flat-top weight x 1)."""
    a = oracle.Sim(_deck([1.0] * S))
    b = oracle.Sim(_no_profile(_deck([1.0] * S)))
    for k in oracle.BEAM_ARRAYS:
        assert np.array_equal(a.get(f"beam1.{k}"), b.get(f"beam1.{k}")), k


@pytest.mark.parametrize("slope", [0.1, -0.03])
def test_profile_weights_closed_form(oracle, slope):
    """Values on a line, v_s = 1 + slope s, at the centres of S equal segments of user xi in [-10, -6] (tail
    first): the weight over the flat-top weight is 1 + slope t, t = (u + 10) / 4 S - 1/2 clamped to [0, S-1].
    A lost centre offset, a reversed orientation or an index off by one fails this."""
    flat = oracle.Sim(_no_profile(_deck([1.0] * S))).get("beam1.w")
    o = oracle.Sim(_deck([1.0 + slope * s for s in range(S)]))
    u = -o.get("beam1.xi")   # user xi of every particle (xi.convention = z-t)
    assert u.min() >= -10 and u.max() <= -6
    t = np.clip((u + 10.0) / 4.0 * S - 0.5, 0, S - 1)
    want = flat * (1.0 + slope * t)
    np.testing.assert_allclose(o.get("beam1.w"), want, rtol=1e-13, atol=0)


def test_profile_must_be_flat_top(oracle):
    deck = _deck([1.0] * S).replace("witness.profile = flattop", "witness.profile = gaussian\nwitness.sigma_xi = 1")
    with pytest.raises(RuntimeError):
        oracle.Sim(deck)


# ------------------------------------------------------------------------------------------ GPU
@pytest.fixture(scope="module")
def qa():
    import quarz_amd
    quarz_amd.lib()
    return quarz_amd


def _same(a, b):
    return np.asarray(a).shape == np.asarray(b).shape and np.array_equal(a, b)


@pytest.mark.gpu
def test_rewind_and_new_profile_equal_oracle(qa, oracle):
    """A run made with profile A, stepped, rewound and given profile B, stepped again: bit-identical to the
    oracle's run made with profile B (fields, beam sources, witness phase space, beam log)."""
    rng = np.random.default_rng(5)
    A, B = rng.uniform(0.5, 1.5, S), rng.uniform(0.5, 1.5, S)
    g = qa.Simulation(_short(_deck(A)))
    g.snapshot()
    g.step()
    g.rewind()
    g.set_beam_profile(1, B)
    g.step()
    o = oracle.Sim(_short(_deck(B)))
    o.step()
    for w in ("fld", "diag", "bsrc"):
        assert _same(getattr(g, w), getattr(o, w)), w
    for bi in range(2):
        for k in qa.BEAM_ARRAYS:
            assert _same(g.get(f"beam{bi}.{k}"), o.get(f"beam{bi}.{k}")), (bi, k)
    da, db = g.beam_diag(1), o.beam_diag(1, order="tree")
    for k in da:
        assert da[k] == db[k], (k, da[k], db[k])


@pytest.mark.gpu
def test_batched_objective_and_fd_gradient(qa, oracle):
    """The workload at full size: J of 2 S + 1 profiles in one batched step.  J(v) equals a plain run's, one
    gradient component equals the oracle's central difference bit for bit (J from moments summed in the
    on-device reduction's order, which the oracle restates), and the gradient predicts the
    central difference along a random direction (at this size J is smooth at the scale of eps: eps 1e-6 and
    1e-7 agree to about 1e-5 relative; a smaller box with 10x fewer beam particles is not)."""
    from quarz_amd import witness
    v = np.array(configs.c4_initial_profile())
    deck = configs.c4_deck(profile=v)
    obj = witness.SpreadObjective(deck, beam=1, nseg=S, method="fd")
    eps = 1e-6
    J0, g = obj.value_and_grad(v, eps)
    plain = qa.Simulation(deck)
    plain.step()
    assert J0 == witness.relative_spread(plain.beam_diag(1))
    s = 13
    Jp = _oracle_J(oracle, deck, v, s, eps)
    Jm = _oracle_J(oracle, deck, v, s, -eps)
    assert g[s] == (Jp - Jm) / (2 * eps), (g[s], (Jp - Jm) / (2 * eps))
    d = np.random.default_rng(9).normal(0, 1, S)
    d /= np.linalg.norm(d)
    Jd = obj.evaluate([v + eps * d, v - eps * d])
    fd = (Jd[0] - Jd[1]) / (2 * eps)
    assert abs(fd - g @ d) <= 1e-3 * abs(g @ d), (fd, g @ d)
    obj.close()


def _oracle_J(oracle, deck, v, s, step):
    from quarz_amd import witness
    vv = np.array(v, dtype=np.float64)
    vv[s] += step
    old = "witness.pwl_weights = " + " ".join(f"{x:.17g}" for x in v)
    assert old in deck
    o = oracle.Sim(deck.replace(old, "witness.pwl_weights = " + " ".join(f"{x:.17g}" for x in vv)))
    o.step()
    return witness.relative_spread(o.beam_diag(1, order="tree"))


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["adjoint", "fd"])
def test_adam_lowers_the_spread(qa, method):
    """A few Adam iterations on the profile, with the adjoint's or the finite-difference gradient behind
    torch.autograd."""
    from quarz_amd import witness
    obj = witness.SpreadObjective(configs.c4_deck(), beam=1, nseg=S, method=method)
    v, hist = witness.optimise(obj, configs.c4_initial_profile(), iters=4, lr=0.05)
    obj.close()
    assert hist[-1] < hist[0], hist
    assert np.all(v >= 0.05)
