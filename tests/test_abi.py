"""CPU checks of the C ABI: the library loads and exports every entry point
the header declares, and the Python binding declares them all.  No compute
calls (there is no GPU here)."""
import ctypes
import os

import pytest

from helpers import HEADER, LIB, declared_symbols


def test_header_declares_the_api():
    syms = declared_symbols()
    for s in ("itnbp_create", "itnbp_net_create", "itnbp_update", "itnbp_expect1",
              "itnbp_rdm2", "itnbp_apply_gates", "itnbp_energy_grad", "itnbp_update_batched",
              "itnbp_logscalar", "itnbp_normalize"):
        assert s in syms


@pytest.mark.skipif(not os.path.exists(LIB), reason="libitnbp.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    import itnbp._lib as L
    assert set(declared_symbols()) == set(L.SIGNATURES)


def test_missing_library_fails_loudly(tmp_path):
    import itnbp._lib as L
    with pytest.raises(L.ITNBPLibraryMissing):
        L.load(str(tmp_path / "nope.so"))
