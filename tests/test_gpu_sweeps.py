"""GPU parity of the level-scheduled sequential sweeps against the oracle: the DIC preconditioner and
Gauss-Seidel to the bit, and OpenFOAM's PCG with DIC -- to the bit with sequential reductions
(OpenFOAM's 224 iterations on the pitzDaily dump), within one iteration with the tree reductions."""
import ctypes
import os

import numpy as np
import pytest

import gpu_util as G
import oracle_lib as O
from brae_amd import hexmesh, mfv

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def pitz():
    d = np.load(os.path.join(GOLD, "pitzdaily.npz"))
    m = O.Mesh.from_npz(d)
    nI = m.nInternal
    diag, up, _ = O.laplacian(m, m.deltaCoeffs, np.ones(nI))
    types = np.zeros(m.nBFaces, np.int32)
    for name, t in zip(m.patchNames, d["T_types"]):
        types[m.patch_slice(name)] = {"fixedValue": 1, "zeroGradient": 0, "empty": 2}[str(t)]
    iC, bC = O.laplacian_bcoeffs(m, types, d["T_values"], np.ones(m.nBFaces))
    diagC, b = O.fold(m, -diag, np.zeros(m.nCells), types, -iC, -bC)
    return "pitzDaily", m.nCells, m.own(), m.neighbour, m.bFaceCell, diagC, -up, b, d


def channel():
    m = hexmesh.box_mesh(30, 14, 9, length=(4.0, 1.0, 0.5), grading=(1.0, 20.0, 20.0), jitter=0.1, seed=9)
    nI, nC = m["nInternal"], m["nCells"]
    own = np.ascontiguousarray(m["owner"][:nI])
    rng = np.random.default_rng(9)
    diag, up = np.zeros(nC), np.zeros(nI)
    O.lib().orc_laplacian(nC, nI, O.p(own), O.p(m["neighbour"]), O.p(m["nonOrthDeltaCoeffs"]),
                          O.p(rng.uniform(0.5, 2.0, nI)), O.p(O.f64(m["magSf"][:nI])), O.p(diag), O.p(up), None)
    diag, up = -diag, -up
    diag[0] *= 2.0
    return "channel", nC, own, m["neighbour"], np.ascontiguousarray(m["owner"][nI:]), diag, up, rng.normal(size=nC), None


@pytest.fixture(scope="module", params=["pitzDaily", "channel"])
def system(request):
    return pitz() if request.param == "pitzDaily" else channel()


def test_dic_and_gauss_seidel_bit_exact(system):
    label, nC, own, nei, bfc, diag, up, b, _ = system
    print("case:", label, flush=True)
    case = G.Case(nC, own, nei, bfc)
    A = mfv.ldu(G.dev(diag), G.dev(up))
    r = np.random.default_rng(1).normal(size=nC)
    dw = G.zeros(nC)
    case.call("mfv_dic_precondition", case.mesh.h, ctypes.byref(A), mfv.ptr(G.dev(r)), mfv.ptr(dw))
    assert np.array_equal(G.host(dw), O.dic_precondition(nC, own, nei, diag, up, r))
    x0 = np.random.default_rng(2).normal(size=nC)
    dx = G.dev(x0)
    case.call("mfv_gauss_seidel", case.mesh.h, ctypes.byref(A), mfv.ptr(G.dev(b)), mfv.ptr(dx), 3)
    case.sync()
    assert np.array_equal(G.host(dx), O.gauss_seidel(nC, own, nei, diag, up, b, x0, 3))


@pytest.mark.parametrize("sequential", [True, False])
def test_pcg_dic(system, sequential):
    label, nC, own, nei, bfc, diag, up, b, d = system
    print("case:", label, "sequential" if sequential else "tree", flush=True)
    case = G.Case(nC, own, nei, bfc)
    if sequential:
        case.ctx.sequential_reduce(True)
    A = mfv.ldu(G.dev(diag), G.dev(up))
    c = mfv.PcgControls()
    mfv.lib().mfv_pcg_default_controls(ctypes.byref(c))
    c.tolerance, c.relTol, c.maxIter, c.checkEvery = 1e-8, 0.0, 1000, 8
    psi = G.zeros(nC)
    perf = mfv.Perf()
    case.call("mfv_pcg_dic_solve", case.mesh.h, ctypes.byref(A), mfv.ptr(G.dev(b)), mfv.ptr(psi), None,
              ctypes.byref(c), ctypes.byref(perf))
    ref_psi, ref = O.pcg_dic(nC, own, nei, diag, up, b, np.zeros(nC), 1e-8)
    x = G.host(psi)
    if sequential:
        assert perf.nIterations == ref.nIterations
        assert perf.initialResidual == ref.initialResidual and perf.finalResidual == ref.finalResidual
        assert np.array_equal(x, ref_psi)
        if d is not None:
            assert perf.nIterations == int(d["pcg_iters"]) == 224   # OpenFOAM's own count on this system
    else:
        assert abs(perf.nIterations - ref.nIterations) <= 1 and perf.finalResidual < 1e-8
        assert np.abs(x - ref_psi).max() <= 1e-6 * np.abs(ref_psi).max()


def test_sym_gauss_seidel_bit_exact(system):
    label, nC, own, nei, bfc, diag, up, b, _ = system
    print("case:", label, "symGaussSeidel", flush=True)
    case = G.Case(nC, own, nei, bfc)
    A = mfv.ldu(G.dev(diag), G.dev(up))
    x0 = np.random.default_rng(3).normal(size=nC)
    dx = G.dev(x0)
    case.call("mfv_sym_gauss_seidel", case.mesh.h, ctypes.byref(A), mfv.ptr(G.dev(b)), mfv.ptr(dx), 2)
    assert np.array_equal(G.host(dx), O.gauss_seidel(nC, own, nei, diag, up, b, x0, 2, symmetric=True))
