"""The C restatement (oracle.c) against the reference's own host C++, compiled from its sources into
oracle/_ref/libbrae_ref.so (`make -C oracle ref`, where the reference tree is mounted). Bit-exact:
both sides are plain FP64 loops in the same order. Skipped where the library was not built.

Covered: Amul (ldu_spmv.cu:5-24), the OpenFOAM PCG with DIC (pcg.cu:6-149), every pass of the
faceAreaPair agglomeration (pair_gamg_agglomeration_cpp.cu:158-284, the level loop's weight
restriction :313-340) and lduAddressing::band (:347-365)."""
import ctypes
import os

import numpy as np
import pytest

import oracle_lib as O
from brae_amd import hexmesh

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                   "libbrae_ref.so")
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="reference host library not built (oracle/_ref)")

p = O.p
i32, f64 = O.i32, O.f64


@pytest.fixture(scope="module")
def ref():
    L = ctypes.CDLL(REF)
    I, D, V = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    sigs = {
        "ref_amul": [I, I] + [V] * 7,
        "ref_pcg": [I, I] + [V] * 5 + [I] + [V] * 4 + [D, D, I, I, V],
        "ref_pair_agglomerate": [I, I, V, V, V, V, V],
        "ref_ldu_band": [I, I, V, V, V, V],
    }
    for n, a in sigs.items():
        getattr(L, n).restype = I
        getattr(L, n).argtypes = a
    return L


def channel(nx, ny, nz, seed):
    m = hexmesh.box_mesh(nx, ny, nz, (4.0, 1.0, 0.5), (1.0, 20.0, 20.0), 0.1, seed=seed)
    nI = m["nInternal"]
    return m, i32(m["owner"][:nI]), i32(m["neighbour"]), i32(m["owner"][nI:])


@pytest.fixture(scope="module")
def pitz():
    d = np.load(os.path.join(GOLD, "pitzdaily.npz"))
    return O.Mesh.from_npz(d)


def ref_amul(L, nC, own, nei, diag, upper, lower, psi):
    out = np.zeros(nC)
    assert L.ref_amul(nC, len(own), p(own), p(nei), p(f64(diag)), p(f64(upper)),
                      None if lower is None else p(f64(lower)), p(f64(psi)), p(out)) == 0
    return out


def test_amul_bit_exact(ref):
    m, own, nei, _ = channel(12, 9, 7, 5)
    rng = np.random.default_rng(3)
    nC, nI = m["nCells"], len(own)
    diag, upper, lower, psi = rng.normal(size=nC), rng.normal(size=nI), rng.normal(size=nI), rng.normal(size=nC)
    for lo in (lower, None):
        a = O.amul(nC, own, nei, diag, upper, lo, psi)
        b = ref_amul(ref, nC, own, nei, diag, upper, lo, psi)
        assert np.array_equal(a, b)


def pressure_system(m, own, nei, bcell, seed):
    """-laplacian(gamma, p) on the jittered channel, fixedValue at the outlet, zeroGradient elsewhere,
    with a random right-hand side: the matrix, the boundary coefficients and the folded system."""
    rng = np.random.default_rng(seed)
    nC, nI = m["nCells"], len(own)
    nB = len(bcell)
    gamma = rng.uniform(0.5, 2.0, nI)
    diag, up = np.zeros(nC), np.zeros(nI)
    O.lib().orc_laplacian(nC, nI, p(own), p(nei), p(f64(m["nonOrthDeltaCoeffs"])), p(gamma),
                          p(f64(m["magSf"][:nI])), p(diag), p(up), None)
    types = np.zeros(nB, np.int32)
    for name, start, size in m["patches"]:
        if name == "outlet":
            types[start - nI:start - nI + size] = O.BC_FIXED_VALUE
    iC, bC = np.zeros(nB), np.zeros(nB)
    O.lib().orc_laplacian_bcoeffs(nB, p(types), p(np.zeros(nB)), p(np.ones(nB)), p(f64(m["magSf"][nI:])),
                                  p(f64(m["bDeltaCoeffs"])), p(iC), p(bC))
    source = rng.normal(size=nC)
    # positive-definite form: negate the laplacian
    diag, up, iC, bC = -diag, -up, -iC, -bC
    diagC, b = np.zeros(nC), np.zeros(nC)
    O.lib().orc_fold(nC, p(diag), p(source), nB, p(bcell), p(types), p(iC), p(bC), p(diagC), p(b))
    return diag, up, source, iC, bC, diagC, b


@pytest.mark.parametrize("tol,relTol", [(1e-9, 0.0), (1e-20, 1e-3)])
def test_pcg_dic_bit_exact(ref, tol, relTol):
    m, own, nei, bcell = channel(10, 8, 6, 11)
    nC = m["nCells"]
    diag, up, source, iC, bC, diagC, b = pressure_system(m, own, nei, bcell, 4)
    psi0 = np.random.default_rng(8).normal(size=nC) * 0.1
    psi_o, perf = O.pcg_dic(nC, own, nei, diagC, up, b, psi0, tol, relTol=relTol, maxIter=1000)
    psi_r = f64(psi0).copy()
    rp = np.zeros(3)
    assert ref.ref_pcg(nC, len(own), p(own), p(nei), p(f64(diag)), p(f64(up)), p(f64(source)), len(bcell),
                       p(bcell), p(f64(iC)), p(f64(bC)), p(psi_r), tol, relTol, 1000, 0, p(rp)) == 0
    assert perf.nIterations == int(rp[2]) and perf.nIterations > 5
    assert perf.initialResidual == rp[0] and perf.finalResidual == rp[1]
    assert np.array_equal(psi_o, psi_r)


def test_pcg_dic_pitzdaily_bit_exact(ref, pitz):
    """the OpenFOAM dump's system (test_pcg.cu:31-38) through both"""
    mesh = pitz
    nC, nI = mesh.nCells, mesh.nInternal
    diag, upper, _ = O.laplacian(mesh, mesh.deltaCoeffs, np.ones(nI))
    b0 = np.random.default_rng(1).normal(size=nC)
    diagC, b = O.fold(mesh, -diag, b0, np.zeros(mesh.nBFaces, np.int32), np.zeros(mesh.nBFaces),
                      np.zeros(mesh.nBFaces))
    diagC[0] *= 2.0   # pin the level (a reference cell)
    psi_o, perf = O.pcg_dic(nC, mesh.own(), mesh.neighbour, diagC, -upper, b, np.zeros(nC), 1e-8)
    psi_r, rp = np.zeros(nC), np.zeros(3)
    zb = np.zeros(mesh.nBFaces)
    assert ref.ref_pcg(nC, nI, p(mesh.own()), p(mesh.neighbour), p(diagC), p(f64(-upper)), p(f64(b)),
                       mesh.nBFaces, p(mesh.bFaceCell), p(zb), p(zb), p(psi_r), 1e-8, 0.0, 1000, 0, p(rp)) == 0
    assert perf.nIterations == int(rp[2])
    assert np.array_equal(psi_o, psi_r)


@pytest.mark.parametrize("shape,seed", [((12, 10, 6), 1), ((25, 7, 9), 2), ((40, 30, 1), 3)])
def test_faceareapair_every_pass_bit_exact(ref, shape, seed):
    m, own, nei, _ = channel(*shape, seed)
    nC, nI = m["nCells"], len(own)
    w = O.faceareapair_weights(m["Sf"][:nI], m["magSf"][:nI])
    ag = O.of_agglomeration(nC, own, nei, w, nCellsInCoarsestLevel=10)
    assert ag.nLevels >= 3
    fwd = ctypes.c_int(1)
    lo, up, n, wl = own, nei, nC, w
    for lvl in ag.levels:
        rmap = np.zeros(n, np.int32)
        nCoarse = ref.ref_pair_agglomerate(n, len(lo), p(i32(lo)), p(i32(up)), p(f64(wl)), ctypes.byref(fwd),
                                           p(rmap))
        assert nCoarse == lvl["nCoarse"]
        assert np.array_equal(rmap, lvl["map"])
        # the next level's addressing (oracle) and weights (reference loop: sum over faceRestrict >= 0,
        # ascending fine face)
        fr = lvl["faceRestrict"]
        wc = np.zeros(lvl["nCoarseFaces"])
        for f in range(len(fr)):
            if fr[f] >= 0:
                wc[fr[f]] += wl[f]
        lo, up, n, wl = lvl["cOwn"], lvl["cNei"], lvl["nCoarse"], wc
        # the coarse addressing is upper-triangular and sorted by owner
        assert np.all(lo < up) and np.all(np.diff(lo) >= 0)
        bw, prof = ctypes.c_int(), ctypes.c_double()
        assert ref.ref_ldu_band(n, len(lo), p(i32(lo)), p(i32(up)), ctypes.byref(bw), ctypes.byref(prof)) == 0
        cb = np.zeros(n, np.int64)
        np.maximum.at(cb, up, up - lo)
        assert bw.value == cb.max() and prof.value == float(cb.sum())
