"""CPU checks of the C ABI: the library loads and exports every symbol the public headers
declare (no compute calls; there is no GPU here)."""
import ctypes
import os

import pytest

from brae_amd import mfv


def test_library_exports_every_declared_symbol():
    if not os.path.exists(mfv.LIB_PATH):
        pytest.fail("libmfv.so not built: run __graft_entry__.build()")
    L = ctypes.CDLL(mfv.LIB_PATH)
    names = mfv.declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_declares_every_symbol():
    names = mfv.declared_symbols()
    lib = mfv.lib()
    for n in names:
        assert n in mfv._SIGS, n
        assert getattr(lib, n).argtypes is not None


def test_context_without_gpu_fails_loudly():
    """No GPU in this container: creating a context must return an error status, not crash."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    rc = mfv.lib().mfv_ctx_create(0, None, ctypes.byref(h))
    assert rc != 0
    assert b"hip" in (mfv.lib().mfv_last_error(None) or b"")
