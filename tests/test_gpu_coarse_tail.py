"""The coarsest grids' three ways of running the cycle give the same bits.

Below 4096 cells the cycle runs from the first such grid down and back in one workgroup (tailK), or in
one workgroup with the coarsest grids' coefficients, columns, lists and vectors staged in LDS (tailMixK,
the default: as many of the coarsest grids as fit, the others read in place), or grid by grid with the
per-grid kernels (mode 0, what the sequential reductions use). Every row sum and every step is the same
arithmetic in the same order, so one V-cycle and a whole AMG-PCG solve (tree reductions) agree to the bit
across the three, in every cycle precision, with and without pre-sweeps and with extra post-sweeps; the
scaled correction keeps tailK. The per-grid path itself is held to the oracle by test_gpu_amg's V-cycle
tests.
"""
import numpy as np
import pytest

import gpu_util as G
from brae_amd import mfv
from test_gpu_amg import channel_system, pitz_system

pytestmark = pytest.mark.gpu

PRECS = [mfv.PREC_FP64, mfv.PREC_FP32, mfv.PREC_FP16]


@pytest.fixture(scope="module")
def systems():
    return [pitz_system(), channel_system(), channel_system(96, 48, 30, seed=9)]


@pytest.mark.parametrize("prec", PRECS, ids=["fp64", "fp32", "fp16"])
def test_vcycle_same_bits_in_every_mode(systems, prec):
    for S in systems:
        amg = mfv.Amg(S.case.mesh, S.magSf, mfv.AGG_BRAE, 64)
        amg.galerkin(S.dd, S.du, prec)
        tail, staged, lds = amg.coarse_mode()
        print("case:", S.label, S.nC, "precision", prec, "grids", amg.nGrids, "tail from grid", tail,
              "staged from grid", staged, "LDS bytes", lds, flush=True)
        assert tail > 0
        # the coarsest grids are a few hundred cells: at least the two coarsest fit
        assert tail <= staged <= amg.nGrids - 2 and lds > 0
        r = G.dev(np.random.default_rng(21).normal(size=S.nC))
        for sweeps in [(1, 1), (2, 1), (0, 1), (1, 2), (1, 0)]:
            out = {}
            for mode in (0, 1, 2):
                amg.coarse_mode(mode)
                z = G.zeros(S.nC)
                amg.vcycle(r, z, *sweeps)
                out[mode] = G.host(z)
            for mode in (1, 2):
                assert np.array_equal(out[mode], out[0]), (S.label, sweeps, mode, np.abs(out[mode] - out[0]).max())
            assert np.isfinite(out[0]).all() and np.abs(out[0]).max() > 0
        amg.coarse_mode(2)
        amg.close()


@pytest.mark.parametrize("prec", PRECS, ids=["fp64", "fp32", "fp16"])
def test_pcg_same_bits_in_every_mode(systems, prec):
    for S in systems[1:]:
        amg = mfv.Amg(S.case.mesh, S.magSf, mfv.AGG_BRAE, 64)
        amg.galerkin(S.dd, S.du, prec)
        print("case: pcg", S.label, S.nC, "precision", prec, amg.coarse_mode(), flush=True)
        out = {}
        for mode in (0, 1, 2):
            amg.coarse_mode(mode)
            psi = G.zeros(S.nC)
            perf = amg.solve(S.dd, S.du, S.db, psi, tolerance=1e-8, precision=prec, flexible=int(prec != mfv.PREC_FP64),
                             nPreSweeps=1, nPostSweeps=1, scaleCorrection=0, checkEvery=4)
            out[mode] = (G.host(psi), perf.nIterations, perf.finalResidual)
        print("iterations", [out[m][1] for m in out], "final", out[2][2], flush=True)
        assert out[2][2] < 1e-8
        for mode in (1, 2):
            assert out[mode][1] == out[0][1] and out[mode][2] == out[0][2], (S.label, mode)
            assert np.array_equal(out[mode][0], out[0][0]), (S.label, mode)
        amg.coarse_mode(2)
        amg.close()


def test_coarse_mode_refusals(systems):
    S = systems[1]
    amg = mfv.Amg(S.case.mesh, S.magSf, mfv.AGG_BRAE, 64)
    for bad in (-2, 3):
        with pytest.raises(mfv.MfvError):
            amg.coarse_mode(bad)
    amg.close()
