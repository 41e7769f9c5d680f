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


def test_amg_signature_fingerprints_inputs():
    """the hierarchy fingerprint (host only): stable for the same inputs, different when the addressing,
    a weight or a control changes"""
    import numpy as np
    from brae_amd import hexmesh
    m = hexmesh.box_mesh(6, 5, 4, geometry=True)
    nI = m["nInternal"]
    own, nei, w = m["owner"][:nI], m["neighbour"], m["magSf"][:nI]
    s0 = mfv.amg_signature(m["nCells"], own, nei, w)
    assert s0 == mfv.amg_signature(m["nCells"], own, nei, w.copy())
    w2 = w.copy()
    w2[7] = np.nextafter(w2[7], 2.0)
    assert mfv.amg_signature(m["nCells"], own, nei, w2) != s0
    nei2 = nei.copy()
    nei2[[3, 4]] = nei2[[4, 3]]
    assert mfv.amg_signature(m["nCells"], own, nei2, w) != s0
    assert mfv.amg_signature(m["nCells"], own, nei, w, mergeLevels=2) != s0
    assert mfv.amg_signature(m["nCells"], own, nei, w, agglomeration=mfv.AGG_FACEAREAPAIR) != s0
    assert mfv.amg_signature(m["nCells"], own, nei, w, nCellsInCoarsestLevel=8) != s0


def test_faceareapair_weights_match_the_oracle():
    """The library's host restatement of faceAreaPair's face weights equals the oracle's to the bit."""
    import numpy as np
    import oracle_lib as O
    rng = np.random.default_rng(4)
    nF = 1000
    Sf = rng.normal(size=(nF, 3))
    magSf = np.linalg.norm(Sf, axis=1)
    w = np.zeros(nF)
    Sf_c = np.ascontiguousarray(Sf.ravel())
    assert mfv.lib().mfv_faceareapair_weights(nF, Sf_c.ctypes.data, magSf.ctypes.data, w.ctypes.data) == mfv.MFV_OK
    assert np.array_equal(w, O.faceareapair_weights(Sf, magSf))
    assert mfv.lib().mfv_faceareapair_weights(-1, None, None, None) == mfv.MFV_ERR_INVALID
