"""GPU parity of the k-epsilon model (mfv_kepsilon_*) and of the turbulent SIMPLE iteration
(mfv_simple_step_kepsilon) against the oracle, which itself reproduces OpenFOAM's kEpsilon::correct() on
pitzDaily (tests/test_oracle_kepsilon.py).

The model's arithmetic is the host's term for term, except the wall functions' pow(k, 1.5) and log, which come
from the device math library and may differ from the host's libm in the last bit. So:
- the turbulent step's U, p and phi (computed before the model corrects) are the oracle's to the bit with
  sequential reductions and the FP64 cycle;
- k, epsilon and nut (cells, and every boundary face) are held to 1e-11 relative, with the solves' iteration
  counts equal.
"""
import os

import numpy as np
import pytest

import gpu_util as G
import oracle_lib as O
import test_oracle_kepsilon as T
from brae_amd import hexmesh, mfv
from brae_amd.mfv import ptr

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KTOL = 1e-11


def rel(a, b):
    m = np.abs(b).max()
    return np.abs(a - b).max() / (m if m > 0 else 1.0)


class Dev:
    """an O.Mesh's geometry on the device"""

    def __init__(self, mesh):
        self.case = G.Case(mesh.nCells, mesh.own(), mesh.neighbour, mesh.bFaceCell)
        self.d = dict(Sf=G.dev(mesh.Sf.ravel()), Cf=G.dev(mesh.Cf.ravel()), magSf=G.dev(mesh.magSf),
                      C=G.dev(mesh.C.ravel()), V=G.dev(mesh.V), w=G.dev(mesh.weights), dc=G.dev(mesh.deltaCoeffs),
                      ndc=G.dev(mesh.nonOrthDeltaCoeffs), cv=G.dev(mesh.corrVec.ravel()),
                      bdc=G.dev(mesh.bDeltaCoeffs))
        d = self.d
        self.geom = mfv.Geometry(*(ptr(d[k]) for k in ("Sf", "Cf", "magSf", "C", "V", "w", "dc", "ndc", "cv", "bdc")))


def ke_controls(relax=0.7, seq_tol=1e-10):
    return mfv.kepsilon_controls(solver=dict(tolerance=seq_tol, relTol=0.0, maxIter=2000), nu=1e-5, relaxEps=relax,
                                 relaxK=relax)


def compare_ke(dev_arrays, host_arrays, names=("k", "kb", "eps", "epsb", "nut", "nutb")):
    for name, dv, hv in zip(names, dev_arrays, host_arrays):
        got = G.host(dv)
        r = rel(got, hv)
        assert r <= KTOL, (name, r)


@pytest.mark.parametrize("sequential", [True, False], ids=["sequential", "tree"])
def test_kepsilon_correct_pitzdaily(sequential):
    d, mesh, U, Ub, uT, kT, kV, eT, eV, k, kb, eps, epsb, nut, nutb, isWall = T.pitz("corr")
    print("case: kepsilon correct pitzDaily sequential", sequential, flush=True)
    y = O.near_wall_dist(mesh, isWall)
    dv = Dev(mesh)
    ke = mfv.KEpsilon(dv.case.mesh, dv.geom, isWall, y, kT, kV, eT, eV, ke_controls())
    phi, phiB = d["corr_phi0_internal"], d["corr_phi0_boundary"]
    dU, dUb = [G.dev(u) for u in U], [G.dev(u) for u in Ub]
    dk, dkb, de, deb, dn, dnb = (G.dev(a) for a in (k, kb, eps, epsb, nut, nutb))
    dv.case.ctx.sequential_reduce(sequential)
    r = ke.correct(dU, dUb, G.dev(phi), G.dev(phiB), dk, dkb, de, deb, dn, dnb)
    dv.case.ctx.sequential_reduce(False)
    cs = O.kepsilon_case(mesh, isWall, y, kT, kV, eT, eV, nu=1e-5, relaxEps=0.7, relaxK=0.7, tol=1e-10,
                         maxIter=2000)
    ro = O.kepsilon_correct(cs, U, Ub, phi, phiB, k, kb, eps, epsb, nut, nutb)
    assert r.nWallCells == ro.nWallCells
    if sequential:
        assert (r.epsIters, r.kIters) == (ro.epsIters, ro.kIters)
    else:
        assert abs(r.epsIters - ro.epsIters) <= 1 and abs(r.kIters - ro.kIters) <= 1
    assert r.epsFinal < 1e-10 and r.kFinal < 1e-10
    compare_ke((dk, dkb, de, deb, dn, dnb), (k, kb, eps, epsb, nut, nutb))
    # the near-wall epsilon is the wall function's value in both
    wall = np.zeros(mesh.nCells, bool)
    wall[mesh.bFaceCell[isWall == 1]] = True
    assert rel(G.host(de)[wall], eps[wall]) <= 1e-13
    ke.close()


def test_turbulent_simple_step_pitzdaily():
    """one turbulent SIMPLE iteration from OpenFOAM's step-282 fields: nuEff = nu + nut, the step, then
    correct(); U, p and phi to the bit, k, epsilon and nut to 1e-11"""
    d, mesh, U, Ub, uT, kT, kV, eT, eV, k, kb, eps, epsb, nut, nutb, isWall = T.pitz("corr")
    s = np.load(os.path.join(GOLD, "simple_step.npz"))
    sizes = mesh.patchSize
    pT = O.bc_codes(np.repeat(s["p_types"], sizes))
    pV = O.f64(s["p_values"])
    uV = O.f64(d["corr_U_values"])
    print("case: turbulent SIMPLE pitzDaily", mesh.nCells, flush=True)
    y = O.near_wall_dist(mesh, isWall)
    pb = O.evaluate_boundary(mesh.bFaceCell, pT, pV, s["p0"])
    st = O.SimpleState(U, Ub, s["p0"], pb, d["corr_phi0_internal"], d["corr_phi0_boundary"])
    dv = Dev(mesh)
    sim = mfv.Simple(dv.case.mesh, dv.geom, uT, np.ascontiguousarray(uV).ravel(), pT, pV,
                     mfv.simple_controls(U=dict(tolerance=1e-10, preconditioner=mfv.PRECOND_DILU),
                                         p=dict(tolerance=1e-10, precision=mfv.PREC_FP64, nPreSweeps=1,
                                                nPostSweeps=1, flexible=0, scaleCorrection=0, checkEvery=3),
                                         mergeLevels=1, nu=1e-5))
    ke = mfv.KEpsilon(dv.case.mesh, dv.geom, isWall, y, kT, kV, eT, eV, ke_controls())
    dU, dUb = [G.dev(u) for u in st.U], [G.dev(u) for u in st.Ub]
    dp, dpb, dphi, dphiB = G.dev(st.p), G.dev(st.pb), G.dev(st.phi), G.dev(st.phiB)
    dk, dkb, de, deb, dn, dnb = (G.dev(a) for a in (k, kb, eps, epsb, nut, nutb))
    dv.case.ctx.sequential_reduce(True)
    r, kr = mfv.simple_step_kepsilon(sim, ke, dU, dUb, dp, dpb, dphi, dphiB, dk, dkb, de, deb, dn, dnb)
    dv.case.ctx.sequential_reduce(False)
    # the oracle's composition: the step with nuEff = nu + nut, then correct() from its U and phi
    ag = O.brae_agglomeration(mesh.nCells, mesh.own(), mesh.neighbour, O.f64(mesh.magSf[:mesh.nInternal]), target=64)
    cs = O.simple_case(mesh, uT, uV.ravel(), pT, pV, nu=1e-5, pSolver=O.P_BRAE_AMGPCG, aggl=ag, amgPre=1,
                       amgPost=1, nuEff=1e-5 + nut, nuEffB=1e-5 + nutb)
    ro = O.simple_step(cs, st)
    kcs = O.kepsilon_case(mesh, isWall, y, kT, kV, eT, eV, nu=1e-5, relaxEps=0.7, relaxK=0.7, tol=1e-10,
                          maxIter=2000)
    kro = O.kepsilon_correct(kcs, st.U, st.Ub, st.phi, st.phiB, k, kb, eps, epsb, nut, nutb)
    assert list(r.UIters) == list(ro.UIters) and r.pIters[0] == ro.pIters[0]
    for name, dvv, hv in zip(("Ux", "Uy", "Uz", "p", "phi", "phiB"), dU + [dp, dphi, dphiB],
                             st.U + [st.p, st.phi, st.phiB]):
        assert np.array_equal(G.host(dvv), hv), name
    assert (kr.epsIters, kr.kIters) == (kro.epsIters, kro.kIters)
    compare_ke((dk, dkb, de, deb, dn, dnb), (k, kb, eps, epsb, nut, nutb))
    sim.close()
    ke.close()


def test_turbulent_channel_iterations():
    """the C2/C3 channel's construction at a small size (graded, jittered, four walls with wall functions,
    inlet k = 0.375, epsilon = 14.855): three turbulent iterations with the bench's options (linearUpwind U,
    corrected laplacian, one non-orthogonal corrector) against the oracle, FP64 cycle and sequential
    reductions; then the mixed-precision (FP16 cycle, tree reductions) path stays within 1e-4 of it"""
    import test_gpu_simple as S
    m, mesh, uT, uV, pT, pV, st0 = S.channel_case(20, 10, 8)
    nI, nC = mesh.nInternal, mesh.nCells
    nB = mesh.nFaces - nI
    isWall = np.zeros(nB, np.int32)
    kT = np.full(nB, O.BC_ZERO_GRADIENT, np.int32)
    kV, eV = np.zeros(nB), np.zeros(nB)
    for name, start, size in m["patches"]:
        sl = slice(start - nI, start - nI + size)
        if name in ("lowerWall", "upperWall", "back", "front"):
            isWall[sl] = 1
        if name == "inlet":
            kT[sl] = O.BC_FIXED_VALUE
            kV[sl], eV[sl] = 0.375, 14.855
    eT = kT.copy()
    y = O.near_wall_dist(mesh, isWall)
    k0 = np.full(nC, 0.375)
    e0 = np.full(nC, 14.855)
    nut0 = 0.09 * k0 * k0 / e0
    kb0 = O.evaluate_boundary(mesh.bFaceCell, kT, kV, k0)
    eb0 = O.evaluate_boundary(mesh.bFaceCell, eT, eV, e0)
    nutb0 = 0.09 * kb0 * kb0 / np.where(eb0 > 0, eb0, 1.0)
    print("case: turbulent channel", nC, flush=True)
    dv = Dev(mesh)

    def controls(prec):
        return mfv.simple_controls(U=dict(tolerance=1e-9, relTol=0.0),
                                   p=dict(tolerance=1e-9, precision=prec, nPreSweeps=1, nPostSweeps=1,
                                          flexible=int(prec != mfv.PREC_FP64), scaleCorrection=0, checkEvery=3),
                                   mergeLevels=1, nu=1e-5, linearUpwind=1, correctedLaplacian=1, nNonOrthCorr=1)

    results = {}
    for tag, prec, seq in (("fp64", mfv.PREC_FP64, True), ("fp16", mfv.PREC_FP16, False)):
        st = st0.copy()
        sim = mfv.Simple(dv.case.mesh, dv.geom, uT, np.ascontiguousarray(uV).ravel(), pT, pV, controls(prec))
        ke = mfv.KEpsilon(dv.case.mesh, dv.geom, isWall, y, kT, kV, eT, eV, ke_controls(seq_tol=1e-9))
        dU, dUb = [G.dev(u) for u in st.U], [G.dev(u) for u in st.Ub]
        dp, dpb, dphi, dphiB = G.dev(st.p), G.dev(st.pb), G.dev(st.phi), G.dev(st.phiB)
        dk, dkb, de, deb, dn, dnb = (G.dev(a) for a in (k0, kb0, e0, eb0, nut0, nutb0))
        dv.case.ctx.sequential_reduce(seq)
        for _ in range(3):
            r, kr = mfv.simple_step_kepsilon(sim, ke, dU, dUb, dp, dpb, dphi, dphiB, dk, dkb, de, deb, dn, dnb)
            assert kr.nWallCells > 0 and r.pFinal < 1e-9
        dv.case.ctx.sequential_reduce(False)
        results[tag] = [G.host(a) for a in dU + [dp, dk, de, dn]]
        sim.close()
        ke.close()
    # the oracle: the same three iterations
    st = st0.copy()
    k, kb, e, eb, nut, nutb = (a.copy() for a in (k0, kb0, e0, eb0, nut0, nutb0))
    ag = O.brae_agglomeration(nC, mesh.own(), mesh.neighbour, O.f64(mesh.magSf[:nI]), target=64)
    kcs = O.kepsilon_case(mesh, isWall, y, kT, kV, eT, eV, nu=1e-5, relaxEps=0.7, relaxK=0.7, tol=1e-9, maxIter=2000)
    for _ in range(3):
        cs = O.simple_case(mesh, uT, uV.ravel(), pT, pV, nu=1e-5, tolU=1e-9, tolP=1e-9, pSolver=O.P_BRAE_AMGPCG,
                           aggl=ag, amgPre=1, amgPost=1, linearUpwind=1, correctedLaplacian=1, nNonOrthCorr=1,
                           nuEff=1e-5 + nut, nuEffB=1e-5 + nutb)
        O.simple_step(cs, st)
        O.kepsilon_correct(kcs, st.U, st.Ub, st.phi, st.phiB, k, kb, e, eb, nut, nutb)
    ref = st.U + [st.p, k, e, nut]
    for name, got, want in zip(("Ux", "Uy", "Uz", "p", "k", "eps", "nut"), results["fp64"], ref):
        assert rel(got, want) <= 1e-9, (name, rel(got, want))
    for name, got, want in zip(("Ux", "Uy", "Uz", "p", "k", "eps", "nut"), results["fp16"], ref):
        assert rel(got, want) <= 1e-4, (name, rel(got, want))
