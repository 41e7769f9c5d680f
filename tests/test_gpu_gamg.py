"""GPU parity of OpenFOAM's GAMG (solver and PCG preconditioner) against the oracle.

With the context's sequential reductions every sum runs in the host loop's order and the level-scheduled
smoothers equal the host sweeps, so the device solve is held to the oracle to the bit: the solution,
both residuals and the iteration count -- on pitzDaily that count is OpenFOAM's own 101 V-cycles to
1e-8. Every smoother the reference runs is covered, with and without pre-sweeps, and a coarsest level
too big for the one-workgroup solve (the multi-launch PCG-DIC with its singularity guard). With the
tree reductions the solve is held to the oracle's iteration count within two and to its solution."""
import numpy as np
import pytest

import gpu_util as G
import oracle_lib as O
from brae_amd import mfv
from test_gpu_amg import channel_system, pitz_system

pytestmark = pytest.mark.gpu

SMOOTHERS = ["GaussSeidel", "symGaussSeidel", "DIC", "DICGaussSeidel"]


def hierarchy(S, nCoarsest=10):
    w = O.faceareapair_weights(S.Sf, S.magSf)
    ag = O.of_agglomeration(S.nC, S.own, S.nei, w, nCellsInCoarsestLevel=nCoarsest)
    amg = mfv.Amg(S.case.mesh, w, mfv.AGG_FACEAREAPAIR, nCoarsest)
    assert amg.nGrids == ag.nLevels + 1
    return ag, amg


@pytest.fixture(scope="module")
def pitz():
    S = pitz_system()
    ag, amg = hierarchy(S)
    return S, ag, amg


@pytest.fixture(scope="module")
def channel():
    S = channel_system()
    ag, amg = hierarchy(S)
    return S, ag, amg


def same(perf, ref, psi, ref_psi):
    assert perf.nIterations == ref.nIterations
    assert perf.initialResidual == ref.initialResidual
    assert perf.finalResidual == ref.finalResidual
    assert np.array_equal(G.host(psi), ref_psi)


def test_gamg_pitzdaily_bit_exact(pitz):
    S, ag, amg = pitz
    print("case: pitzDaily gamg sequential", flush=True)
    S.case.ctx.sequential_reduce(True)
    try:
        psi = G.zeros(S.nC)
        perf = amg.gamg_solve(S.dd, S.du, S.db, psi, tolerance=1e-8, relTol=0.0, smoother="GaussSeidel")
    finally:
        S.case.ctx.sequential_reduce(False)
    ref_psi, ref = O.gamg_solve(ag, S.diag, S.upper, S.b, np.zeros(S.nC), tolerance=1e-8, relTol=0.0,
                                smoother="GaussSeidel")
    same(perf, ref, psi, ref_psi)
    assert perf.nIterations == 101 and perf.converged   # OpenFOAM's own count on this system


def test_pcg_gamg_pitzdaily_bit_exact(pitz):
    S, ag, amg = pitz
    print("case: pitzDaily pcg-gamg sequential", flush=True)
    S.case.ctx.sequential_reduce(True)
    try:
        psi = G.zeros(S.nC)
        perf = amg.pcg_gamg_solve(S.dd, S.du, S.db, psi, 1e-8, smoother="GaussSeidel")
    finally:
        S.case.ctx.sequential_reduce(False)
    ref_psi, ref = O.pcg_gamg_solve(ag, S.diag, S.upper, S.b, np.zeros(S.nC), 1e-8, smoother="GaussSeidel")
    same(perf, ref, psi, ref_psi)
    assert perf.converged and perf.nIterations < 40


@pytest.mark.parametrize("smoother", SMOOTHERS)
@pytest.mark.parametrize("nPre", [0, 1])
def test_gamg_smoothers_bit_exact(channel, smoother, nPre):
    S, ag, amg = channel
    print("case: channel gamg", smoother, "nPre", nPre, flush=True)
    kw = dict(tolerance=1e-9, relTol=0.0, maxIter=12, smoother=smoother, nPreSweeps=nPre)
    S.case.ctx.sequential_reduce(True)
    try:
        psi = G.zeros(S.nC)
        perf = amg.gamg_solve(S.dd, S.du, S.db, psi, **kw)
    finally:
        S.case.ctx.sequential_reduce(False)
    ref_psi, ref = O.gamg_solve(ag, S.diag, S.upper, S.b, np.zeros(S.nC), **kw)
    same(perf, ref, psi, ref_psi)
    assert perf.finalResidual < 1e-2 * perf.initialResidual


@pytest.mark.parametrize("smoother,nV,scale", [("DIC", 1, 1), ("symGaussSeidel", 2, 0), ("DICGaussSeidel", 3, 1)])
def test_pcg_gamg_variants_bit_exact(channel, smoother, nV, scale):
    S, ag, amg = channel
    print("case: channel pcg-gamg", smoother, nV, scale, flush=True)
    kw = dict(smoother=smoother, nVcycles=nV, scaleCorrection=scale, tolerance=1e-3)   # coarsest solve's
    S.case.ctx.sequential_reduce(True)
    try:
        psi = G.zeros(S.nC)
        perf = amg.pcg_gamg_solve(S.dd, S.du, S.db, psi, 1e-9, 0.0, 200, 0, **kw)
    finally:
        S.case.ctx.sequential_reduce(False)
    ref_psi, ref = O.pcg_gamg_solve(ag, S.diag, S.upper, S.b, np.zeros(S.nC), 1e-9, 0.0, 200, 0, **kw)
    same(perf, ref, psi, ref_psi)
    assert perf.converged


def test_gamg_large_coarsest_level():
    """A coarsest level too big for LDS: the multi-launch guarded PCG-DIC, still to the bit."""
    S = channel_system()
    ag, amg = hierarchy(S, nCoarsest=1500)
    n, f = amg.grid_size(amg.nGrids - 1)
    print("case: channel gamg coarsest", n, "cells", f, "faces", flush=True)
    assert 64 * n + 24 * f > 60 * 1024
    S.case.ctx.sequential_reduce(True)
    psi = G.zeros(S.nC)
    perf = amg.gamg_solve(S.dd, S.du, S.db, psi, tolerance=1e-9, maxIter=6)
    ref_psi, ref = O.gamg_solve(ag, S.diag, S.upper, S.b, np.zeros(S.nC), tolerance=1e-9, maxIter=6)
    same(perf, ref, psi, ref_psi)


def test_gamg_tree_reductions_converge_like_oracle(pitz):
    S, ag, amg = pitz
    print("case: pitzDaily gamg tree", flush=True)
    psi = G.zeros(S.nC)
    perf = amg.gamg_solve(S.dd, S.du, S.db, psi, tolerance=1e-8, relTol=0.0)
    ref_psi, ref = O.gamg_solve(ag, S.diag, S.upper, S.b, np.zeros(S.nC), tolerance=1e-8, relTol=0.0)
    assert perf.converged and perf.finalResidual < 1e-8
    assert abs(perf.nIterations - ref.nIterations) <= 2
    x = G.host(psi)
    assert np.abs(x - ref_psi).max() <= 1e-6 * np.abs(ref_psi).max()
    psi = G.zeros(S.nC)
    perf = amg.pcg_gamg_solve(S.dd, S.du, S.db, psi, 1e-8)
    ref_psi, ref = O.pcg_gamg_solve(ag, S.diag, S.upper, S.b, np.zeros(S.nC), 1e-8)
    assert perf.converged and abs(perf.nIterations - ref.nIterations) <= 1
    assert np.abs(G.host(psi) - ref_psi).max() <= 1e-6 * np.abs(ref_psi).max()


def test_gamg_refusals(channel):
    S, ag, amg = channel
    psi = G.zeros(S.nC)
    with pytest.raises(mfv.MfvError, match="smoother"):
        amg.gamg_solve(S.dd, S.du, S.db, psi, smoother=7)
    with pytest.raises(mfv.MfvError, match="nVcycles"):
        amg.pcg_gamg_solve(S.dd, S.du, S.db, psi, 1e-6, nVcycles=0)
    lower = G.dev(S.upper * 1.5)
    A = mfv.ldu(S.dd, S.du, lower)
    import ctypes
    c = mfv.gamg_controls()
    with pytest.raises(mfv.MfvError, match="asymmetric"):
        S.case.call("mfv_gamg_solve", S.case.mesh.h, amg.h, ctypes.byref(A), mfv.ptr(S.db), mfv.ptr(psi),
                    ctypes.byref(c), None)
