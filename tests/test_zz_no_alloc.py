"""qz_sweep_noalloc, the sweep entry point that allocates nothing (its argument block and workspace live in the
library's static device array), against qz_sweep on the same prepared runs, through the C-ABI: bit-identical
fields, diagnostics and plasma state, on two calls in a row, the second with a smaller run (fewer nodes and
slices) that reuses the array the first one left."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DECK_A = {"modes": 1, "xi.min": -5, "grid.dr": 0.04, "driver.n0": 0.5, "driver.x0": 0.05,
          "driver.nparticles": 20000, "driver.sigma_r": 0.5, "driver.rigid": 0, "electrons.ppc": 1}
DECK_B = dict(DECK_A, **{"xi.min": -3, "grid.dr": 0.08})   # 101 nodes and 301 slices against 201 and 501


@pytest.fixture(scope="module")
def qa():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import quarz_amd
    quarz_amd.lib()
    return quarz_amd


def _sweep(qa, deck, ov, entry):
    """One sweep of a prepared run through the raw C-ABI entry point `entry`; the run's arrays afterwards."""
    import torch
    L = qa.lib()
    s = qa.Simulation(deck, ov)
    args = C.create_string_buffer(4096)          # larger than qz_sweep_args
    qa.check(L.qz_sim_prepare_sweep(s.h, args))  # beam deposit, plasma load, background; the run's stream is done
    qa.check(getattr(L, entry)(args, None))
    torch.cuda.synchronize()
    out = {w: getattr(s, w) for w in ("fld", "diag")}
    for k in ("x", "y", "px", "py", "dl", "w"):
        out[f"sp0.{k}"] = s.get(f"sp0.{k}")
    return out, s


def test_noalloc_sweep_equals_sweep(qa, decks):
    deck = f"{decks}/linear.in"
    for ov in (DECK_A, DECK_B):   # the second call is smaller and follows the first on the same scratch array
        want, s1 = _sweep(qa, deck, ov, "qz_sweep")
        got, s2 = _sweep(qa, deck, ov, "qz_sweep_noalloc")
        assert s1.M == s2.M and s1.K == s2.K
        for name in want:
            assert np.array_equal(got[name], want[name]), (ov["grid.dr"], name)
