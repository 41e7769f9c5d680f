"""Small AMG-PCG solves batched into one launch (mfv_amg_pcg_solve_batch), one workgroup per solve.

- A solve's result is the same bits in any batch: each system solved alone in a batch of one against
  the same system in a batch of several (different sizes, seeds and warm starts), in FP64, FP32 and FP16.
- It is the solve of mfv_amg_pcg_solve (the per-grid kernels) up to the order of the sums: both reach
  the tolerance on OpenFOAM's normalised residual, in iteration counts within two of each other, with
  solutions that agree to the tolerance's reach; the initial residuals (no cycle yet) agree to rounding.
- With the scaled correction (V(2,1), the serving loop's cycle) likewise against the per-grid solve.
- Refusals: a hierarchy twice in one batch, Galerkin values in another precision.
"""
import time

import numpy as np
import pytest

import gpu_util as G
from brae_amd import mfv
from test_gpu_amg import channel_system

pytestmark = pytest.mark.gpu

PRECS = [mfv.PREC_FP64, mfv.PREC_FP32, mfv.PREC_FP16]
CTL = dict(tolerance=1e-8, flexible=1, nPreSweeps=1, nPostSweeps=1, scaleCorrection=0)


@pytest.fixture(scope="module")
def systems():
    shapes = [(40, 20, 12, 5), (33, 17, 9, 6), (24, 24, 24, 7), (60, 30, 10, 8), (20, 10, 6, 9)]
    return [channel_system(*s) for s in shapes]


@pytest.mark.parametrize("prec", PRECS, ids=["fp64", "fp32", "fp16"])
def test_batch_same_bits_alone_and_together(systems, prec):
    # the five systems' meshes and hierarchies on one context, so that one call batches them all
    ctx = mfv.Context(0)
    try:
        meshes, amgs = [], []
        for S in systems:
            m = mfv.Mesh(ctx, S.nC, S.own, S.nei, S.bfc)
            a = mfv.Amg(m, S.magSf, mfv.AGG_BRAE, 64)
            a.galerkin(S.dd, S.du, prec)
            meshes.append(m)
            amgs.append(a)
        rng = np.random.default_rng(3)
        warm = [rng.normal(size=S.nC) * 1e-3 for S in systems]
        together = {}
        jobs, psis = [], []
        for k, S in enumerate(systems):
            psi = G.dev(warm[k])
            psis.append(psi)
            jobs.append(amgs[k].solve_batch_job(S.dd, S.db, psi, precision=prec, **CTL))
        t = time.perf_counter()
        perfs = mfv.solve_batch(ctx, jobs)
        dt = time.perf_counter() - t
        print("case: batch of", len(jobs), "precision", prec, "iterations", [p.nIterations for p in perfs],
              "%.1f ms" % (dt * 1e3), flush=True)
        for k in range(len(systems)):
            together[k] = (G.host(psis[k]), perfs[k])
        for k, S in enumerate(systems):
            psi = G.dev(warm[k])
            p1 = mfv.solve_batch(ctx, [amgs[k].solve_batch_job(S.dd, S.db, psi, precision=prec, **CTL)])[0]
            x1 = G.host(psi)
            xt, pt = together[k]
            assert p1.nIterations == pt.nIterations and p1.finalResidual == pt.finalResidual, k
            assert p1.initialResidual == pt.initialResidual
            assert np.array_equal(x1, xt), k
            assert pt.converged and pt.finalResidual < 1e-8
        for a in amgs:
            a.close()
        for m in meshes:
            m.close()
    finally:
        ctx.close()


@pytest.mark.parametrize("prec", PRECS, ids=["fp64", "fp32", "fp16"])
@pytest.mark.parametrize("scaled", [0, 1], ids=["plain", "scaled"])
def test_batch_matches_the_per_grid_solve(systems, prec, scaled):
    ctl = dict(CTL, scaleCorrection=scaled, nPreSweeps=2 if scaled else 1)
    for S in systems:
        amg = mfv.Amg(S.case.mesh, S.magSf, mfv.AGG_BRAE, 64)
        amg.galerkin(S.dd, S.du, prec)
        psi_b = G.zeros(S.nC)
        pb = mfv.solve_batch(S.case.ctx, [amg.solve_batch_job(S.dd, S.db, psi_b, precision=prec, **ctl)])[0]
        psi_k = G.zeros(S.nC)
        pk = amg.solve(S.dd, S.du, S.db, psi_k, precision=prec, checkEvery=4, **ctl)
        print("case:", S.label, S.nC, "precision", prec, "scaled", scaled, "iterations batch/per-grid", pb.nIterations,
              pk.nIterations, "final", pb.finalResidual, pk.finalResidual, flush=True)
        assert pb.converged and pb.finalResidual < 1e-8 and pk.finalResidual < 1e-8
        assert abs(pb.nIterations - pk.nIterations) <= 2
        assert abs(pb.initialResidual - pk.initialResidual) <= 1e-12 * pk.initialResidual
        xb, xk = G.host(psi_b), G.host(psi_k)
        assert np.abs(xb - xk).max() <= 1e-5 * np.abs(xk).max()
        amg.close()


def test_batch_refusals(systems):
    S = systems[0]
    amg = mfv.Amg(S.case.mesh, S.magSf, mfv.AGG_BRAE, 64)
    amg.galerkin(S.dd, S.du, mfv.PREC_FP32)
    psi = G.zeros(S.nC)
    job = amg.solve_batch_job(S.dd, S.db, psi, precision=mfv.PREC_FP32, **CTL)
    with pytest.raises(mfv.MfvError) as e:
        mfv.solve_batch(S.case.ctx, [job, job])
    assert "twice" in str(e.value)
    with pytest.raises(mfv.MfvError):   # the Galerkin values are FP32, not FP16
        mfv.solve_batch(S.case.ctx, [amg.solve_batch_job(S.dd, S.db, psi, precision=mfv.PREC_FP16, **CTL)])
    amg.close()
