"""The xi-pipeline message (qz_pack_message / qz_unpack_message, the device-resident restatement of
Simulation::make_message, src/Simulation.cpp:851-870, and PlasmaSpecies::pack_state / unpack_state,
src/PlasmaSpecies.cpp:317-364), through the C-ABI.

A sweep split at slice k and handed over through the message continues bit for bit: run A sweeps [0, k), packs
its state; run B, prepared from the same deck, unpacks it and sweeps [k, K); the fields and diagnostics of A's
slices and of B's slices, and B's plasma state at the end, equal one uninterrupted sweep of run C.  The message's
layout is checked against run A's own arrays: header, per-species particle arrays, counters, and the guard row
(the beam sources of slice k - 1, padded to 15 components)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BASE = {"xi.min": -5, "grid.dr": 0.04, "driver.n0": 0.5, "driver.x0": 0.05, "driver.nparticles": 20000,
        "driver.sigma_r": 0.5, "driver.rigid": 0, "electrons.ppc": 1}
CASES = {
    "m0": dict(BASE, **{"modes": 0, "driver.x0": 0}),
    "m1": dict(BASE, modes=1),
    "m1_two_species": dict(BASE, **{"modes": 1, "plasma.species": "electrons ions", "ions.ppc": 1,
                                    "ions.charge": 1, "ions.mass": 1836.15, "ions.mobile": 1}),
}
SP_ARRAYS = ("x", "y", "px", "py", "dl", "w")


@pytest.fixture(scope="module")
def qa():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import quarz_amd
    quarz_amd.lib()
    return quarz_amd


def _prepared(qa, deck, ov):
    L = qa.lib()
    s = qa.Simulation(deck, ov)
    assert L.qz_sweep_args_bytes() <= 4096
    args = C.create_string_buffer(4096)
    qa.check(L.qz_sim_prepare_sweep(s.h, args))
    return s, args


def _launch(qa, args, kb, ke):
    import torch
    head = qa.SweepArgsHeadC.from_buffer(args)
    head.k_begin, head.k_end = kb, ke
    qa.check(qa.lib().qz_sweep(args, None))
    torch.cuda.synchronize()


def _state(s):
    return {f"sp{i}.{k}": s.get(f"sp{i}.{k}") for i in range(s.nspecies) for k in SP_ARRAYS}


@pytest.mark.parametrize("case", sorted(CASES))
def test_split_sweep_through_the_message(qa, decks, case):
    import torch
    L = qa.lib()
    deck, ov = f"{decks}/linear.in", CASES[case]
    c, cargs = _prepared(qa, deck, ov)
    _launch(qa, cargs, 0, 0)
    want_fld, want_diag, want_state = c.fld, c.diag, _state(c)
    K, M, nb = c.K, c.M, c.nbeams
    order = 4   # pusher.ab_order of the deck
    for k in (1, 7, K // 2):   # after the first slice; mid-history (7 % 4 = 3); half way
        a, aargs = _prepared(qa, deck, ov)
        _launch(qa, aargs, 0, k)
        n = L.qz_message_doubles(aargs, nb)
        msg = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
        qa.check(L.qz_pack_message(aargs, nb, 3, 2.5, 0, k, C.c_void_p(msg.data_ptr()), None))
        torch.cuda.synchronize()
        host = msg.cpu().numpy()
        assert not np.isnan(host).any(), "message has unwritten doubles"

        # ---- layout against run A's own arrays
        assert host[0] == 3 and host[1] == 2.5 and host[2] == 0 and not host[3:3 + nb].any()
        off = 3 + nb
        state_a = _state(a)
        for i in range(a.nspecies):
            np_i = state_a[f"sp{i}.x"].size
            assert host[off] == np_i
            off += 1
            for key in SP_ARRAYS:
                assert np.array_equal(host[off:off + np_i], state_a[f"sp{i}.{key}"]), (case, k, i, key)
                off += np_i
            assert np.all(host[off:off + np_i] == 0.0)   # no particle is fresh past the first slice
            off += np_i + 5 * order * np_i
            dxi = host[off + 1] / k   # the state's xi is k slices from the head
            assert host[off] == k and dxi > 0
            for sl in range(order):   # the xi of the last push that wrote each history slot
                kw = max([j for j in range(k) if j % order == sl], default=None)
                if kw is None:
                    assert host[off + 2 + sl] == 0.0
                else:
                    assert host[off + 2 + sl] == pytest.approx(dxi * kw, rel=1e-14), (case, k, sl)
            off += 2 + order
        off += 6 * M
        guard = host[off:off + 15 * M].reshape(M, 15)
        assert off + 15 * M == n
        bs = a.bsrc[k - 1]
        assert np.array_equal(guard[:, :bs.shape[1]], bs) and not guard[:, bs.shape[1]:].any()

        # ---- the continuation
        b, bargs = _prepared(qa, deck, ov)
        gdev = torch.zeros(15 * M, dtype=torch.float64, device="cuda")
        qa.check(L.qz_unpack_message(bargs, nb, C.c_void_p(msg.data_ptr()), C.c_void_p(gdev.data_ptr()), None))
        _launch(qa, bargs, k, 0)
        assert np.array_equal(gdev.cpu().numpy(), host[off:off + 15 * M])
        af, bf, ad, bd = a.fld, b.fld, a.diag, b.diag
        assert np.array_equal(af[:k], want_fld[:k]) and np.array_equal(ad[:k], want_diag[:k]), (case, k)
        assert np.array_equal(bf[k:], want_fld[k:]), (case, k, "fields after the hand-over")
        assert np.array_equal(bd[k:], want_diag[k:]), (case, k, "diagnostics after the hand-over")
        got = _state(b)
        for name in want_state:
            assert np.array_equal(got[name], want_state[name]), (case, k, name)
        del a, b
