"""The no-allocation PBiCGStab (mfv_pbicgstab_solve_noalloc) against mfv_pbicgstab_solve, through the C ABI.

The variant runs the same kernels in a static device workspace instead of the mesh's buffers, so on the same
inputs it must give the same bits: psi, iteration count and both residuals. Two calls in a row, the second
on a smaller mesh, so the second runs over the first call's leftover workspace contents; with DILU and the
diagonal preconditioner, tree and sequential reductions. Sizes beyond the workspace are refused, and DILU
on a mesh without a level schedule is refused, both before any launch.
"""
import ctypes

import numpy as np
import pytest

import gpu_util as G
import oracle_lib as O
from brae_amd import hexmesh, mfv
from brae_amd.mfv import ptr

pytestmark = pytest.mark.gpu


def momentum_system(nx, ny, nz, seed):
    """an upwind-convection minus diffusion matrix on a graded, jittered box, with boundary coefficients"""
    m = hexmesh.box_mesh(nx, ny, nz, (4.0, 1.0, 0.5), (1.0, 6.0, 6.0), 0.1, seed=seed)
    nI, nC = m["nInternal"], m["nCells"]
    own, nei, bfc = O.i32(m["owner"][:nI]), O.i32(m["neighbour"]), O.i32(m["owner"][nI:])
    rng = np.random.default_rng(seed)
    phi = rng.normal(size=nI) * m["magSf"][:nI]
    D, U, L = np.zeros(nC), np.zeros(nI), np.zeros(nI)
    O.lib().orc_div_upwind(nC, nI, O.p(own), O.p(nei), O.p(O.f64(phi)), O.p(D), O.p(U), O.p(L))
    lD, lU = np.zeros(nC), np.zeros(nI)
    O.lib().orc_laplacian(nC, nI, O.p(own), O.p(nei), O.p(O.f64(m["deltaCoeffs"])), O.p(np.full(nI, 0.05)),
                          O.p(O.f64(m["magSf"][:nI])), O.p(lD), O.p(lU), None)
    D = D - lD + 0.2 * np.abs(U).max()
    U, L = U - lU, L - lU
    iC = rng.uniform(0.0, 0.5, len(bfc))
    bC = rng.normal(size=len(bfc))
    src = rng.normal(size=nC) * m["V"]
    return m, own, nei, bfc, D, U, L, src, iC, bC


def solve(case, entry, arrays, x0, controls):
    D, U, L, src, iC, bC = arrays
    dD, dU, dL, ds, dic, dbc, dx = (G.dev(a) for a in (D, U, L, src, iC, bC, x0))
    A = mfv.ldu(dD, dU, dL)
    perf = mfv.Perf()
    case.call(entry, case.mesh.h, ctypes.byref(A), ptr(ds), ptr(dic), ptr(dbc), ptr(dx), ctypes.byref(controls),
              ctypes.byref(perf))
    case.sync()
    return G.host(dx), perf


@pytest.mark.parametrize("sequential", [False, True], ids=["tree", "sequential"])
def test_noalloc_matches_the_allocating_solve(sequential):
    ctx = mfv.Context(0)
    ctx.sequential_reduce(sequential)
    # a larger mesh first, then a smaller one: the second call runs over the first's workspace contents
    for shape in ((40, 30, 20, 7), (13, 9, 5, 8)):
        m, own, nei, bfc, D, U, L, src, iC, bC = momentum_system(*shape)
        nC = m["nCells"]
        print("case: pbicgstab noalloc", shape[:3], "cells", nC, "sequential", sequential, flush=True)
        case = G.Case(nC, own, nei, bfc, ctx=ctx)
        x0 = np.random.default_rng(shape[3]).normal(size=nC)
        for pre in (mfv.PRECOND_DILU, mfv.PRECOND_DIAGONAL):
            c = mfv.bicg_controls(tolerance=1e-12, relTol=0.0, maxIter=400, preconditioner=pre, checkEvery=3)
            x_ref, p_ref = solve(case, "mfv_pbicgstab_solve", (D, U, L, src, iC, bC), x0, c)
            x_na, p_na = solve(case, "mfv_pbicgstab_solve_noalloc", (D, U, L, src, iC, bC), x0, c)
            assert p_ref.nIterations > 0
            assert p_na.nIterations == p_ref.nIterations, (pre, p_na.nIterations, p_ref.nIterations)
            assert p_na.initialResidual == p_ref.initialResidual
            assert p_na.finalResidual == p_ref.finalResidual
            assert p_na.converged == p_ref.converged
            assert np.array_equal(x_na, x_ref), pre
    ctx.sequential_reduce(False)


def test_noalloc_refusals():
    ctx = mfv.Context(0)
    # DILU on a mesh whose level schedule was never built
    m, own, nei, bfc, D, U, L, src, iC, bC = momentum_system(6, 5, 4, 11)
    case = G.Case(m["nCells"], own, nei, bfc, ctx=ctx)
    c = mfv.bicg_controls(tolerance=1e-12, preconditioner=mfv.PRECOND_DILU)
    with pytest.raises(mfv.MfvError, match="level schedule"):
        solve(case, "mfv_pbicgstab_solve_noalloc", (D, U, L, src, iC, bC), np.zeros(m["nCells"]), c)
    # the diagonal preconditioner needs nothing beforehand
    x, p = solve(case, "mfv_pbicgstab_solve_noalloc", (D, U, L, src, iC, bC), np.zeros(m["nCells"]),
                 mfv.bicg_controls(tolerance=1e-12, preconditioner=mfv.PRECOND_DIAGONAL))
    assert p.converged
    # a mesh beyond the static workspace: refused on the host, before any launch
    nx, ny, nz = 101, 81, 65   # 531,765 cells > 524,288
    assert nx * ny * nz > 524288   # MFV_STATIC_BICG_MAX_CELLS
    big = hexmesh.box_mesh(nx, ny, nz, (1.0, 1.0, 1.0), (1.0, 1.0, 1.0), 0.0, seed=0)
    nI, nC = big["nInternal"], big["nCells"]
    bcase = G.Case(nC, big["owner"][:nI], big["neighbour"], big["owner"][nI:], ctx=ctx)
    nB = len(big["owner"]) - nI
    with pytest.raises(mfv.MfvError, match="exceed the static workspace"):
        solve(bcase, "mfv_pbicgstab_solve_noalloc",
              (np.ones(nC), np.zeros(nI), np.zeros(nI), np.zeros(nC), np.zeros(nB), np.zeros(nB)), np.zeros(nC),
              mfv.bicg_controls(preconditioner=mfv.PRECOND_DIAGONAL))
