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
