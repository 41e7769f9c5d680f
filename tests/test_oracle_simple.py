"""The oracle's SIMPLE iteration and its two linear solvers, pinned. CPU only.

- The composed step against OpenFOAM's own dumpSimpleStep output (validation/matrixDumpSimple/step.dat):
  p, U and phi, internal and every patch, at the reference's own 1e-9 gate
  (tests/test_simple_step_cpp.cu:36-40), from OpenFOAM's step-282 fields with phi as read, laminar
  nu = 1e-5, upwind, orthogonal laplacian, U relaxed 0.7 and p 0.3, GAMG for p and PBiCGStab-DILU for U,
  both to 1e-10 as cpu::StepInput defaults them. A control checks the step moved p.
- PBiCGStab-DILU (pbicgstab.cu:6-150) bit for bit against the reference's own compiled pbicgstab.cu.
- The GAMG of gamg.cu:267-400 part by part: each pairing pass bit for bit against the reference's compiled
  faceAreaPair pass (the same walk); the coarsening against the dense Galerkin product R A R^T on integer
  data (exact), with its faces the sorted distinct pairs; GaussSeidel against the closed form
  (D + L)^-1 (b - U x) on an asymmetric matrix; the dense solve against numpy; the scaling against its
  definition; and the whole solve against a dense solve.
- fvc::gradUBoundary against a linear velocity field on a uniform box: the fixedValue faces give the
  exact (non-symmetric) gradient, the zeroGradient faces the cell gradient with its normal row removed.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle_lib as O
from brae_amd import hexmesh

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                   "libbrae_ref.so")
p, i32, f64 = O.p, O.i32, O.f64


def rel(a, b):
    m = np.abs(b).max()
    return np.abs(a - b).max() / (m if m > 0 else 1.0)


@pytest.fixture(scope="module")
def pitz_step():
    mesh = O.Mesh.from_npz(np.load(os.path.join(GOLD, "pitzdaily.npz")))
    d = np.load(os.path.join(GOLD, "simple_step.npz"))
    return mesh, d


def pitz_state(mesh, d):
    sizes = mesh.patchSize
    uT = O.bc_codes(np.repeat(d["U_types"], sizes))
    pT = O.bc_codes(np.repeat(d["p_types"], sizes))
    uV = f64(d["U_values"])
    pV = f64(d["p_values"])
    bfc = mesh.bFaceCell
    U0 = f64(d["U0"])
    U = [U0[:, k].copy() for k in range(3)]
    Ub = [O.evaluate_boundary(bfc, uT, uV[:, k], U[k]) for k in range(3)]
    pb = O.evaluate_boundary(bfc, pT, pV, d["p0"])
    st = O.SimpleState(U, Ub, d["p0"], pb, d["phi0_internal"], d["phi0_boundary"])
    return st, uT, uV, pT, pV


def test_simple_step_matches_openfoam_step_dat(pitz_step):
    mesh, d = pitz_step
    st, uT, uV, pT, pV = pitz_state(mesh, d)
    p0 = st.p.copy()
    cs = O.simple_case(mesh, uT, uV, pT, pV, nu=float(d["nu"]), relaxU=0.7, relaxP=0.3)
    r = O.simple_step(cs, st)
    assert rel(st.p, d["step_p"]) <= 1e-9
    U = np.stack(st.U, 1)
    assert rel(U, d["step_U"]) <= 1e-9
    assert rel(st.phi, d["step_phi_internal"]) <= 1e-9
    for name in mesh.patchNames:
        ref = d["step_phi_" + name]
        if ref.size == 0:
            continue
        got = st.phiB[mesh.patch_slice(name)]
        m = np.abs(ref).max()
        assert np.abs(got - ref).max() <= 1e-9 * (m if m > 0 else 1.0), name
    # control: the step moved the fields (a no-op driver would match nothing above, but say so)
    assert np.abs(st.p - p0).max() > 1.0
    assert r.pIters[0] > 0 and min(r.UIters) > 0
    assert r.pFinal < 1e-10 and max(r.UFinal) < 1e-10


def test_simple_step_solver_variants_agree(pitz_step):
    """the same step with the pressure solved by brae's AMG-PCG (FP64 cycle) or OpenFOAM's PCG + GAMG to the
    same 1e-10 reaches step.dat at the same gate: the composition, not the solver, fixes the answer"""
    mesh, d = pitz_step
    for solver in (O.P_BRAE_AMGPCG, O.P_PCG_GAMG):
        st, uT, uV, pT, pV = pitz_state(mesh, d)
        own, nei = mesh.own(), i32(mesh.neighbour)
        if solver == O.P_BRAE_AMGPCG:
            ag = O.brae_agglomeration(mesh.nCells, own, nei, f64(mesh.magSf[:mesh.nInternal]), target=64)
        else:
            w = O.faceareapair_weights(mesh.Sf[:mesh.nInternal], mesh.magSf[:mesh.nInternal])
            ag = O.of_agglomeration(mesh.nCells, own, nei, w, nCellsInCoarsestLevel=10)
        cs = O.simple_case(mesh, uT, uV, pT, pV, nu=float(d["nu"]), pSolver=solver, aggl=ag)
        O.simple_step(cs, st)
        assert rel(st.p, d["step_p"]) <= 1e-9, solver
        assert rel(np.stack(st.U, 1), d["step_U"]) <= 1e-9, solver
        assert rel(st.phi, d["step_phi_internal"]) <= 1e-9, solver


def test_simple_step_skipped_momentum_predictor_differs(pitz_step):
    """control for the gate: leaving out the momentum solve changes U beyond the gate"""
    mesh, d = pitz_step
    st, uT, uV, pT, pV = pitz_state(mesh, d)
    cs = O.simple_case(mesh, uT, uV, pT, pV, nu=float(d["nu"]), momentumPredictor=0)
    O.simple_step(cs, st)
    assert rel(np.stack(st.U, 1), d["step_U"]) > 1e-6


def channel(nx, ny, nz, seed, jitter=0.1):
    m = hexmesh.box_mesh(nx, ny, nz, (4.0, 1.0, 0.5), (1.0, 6.0, 6.0), jitter, seed=seed)
    nI = m["nInternal"]
    return m, i32(m["owner"][:nI]), i32(m["neighbour"]), i32(m["owner"][nI:])


def momentum_system(m, own, nei, bfc, seed):
    """upwind convection of a random flux minus a laplacian, with boundary coefficients: asymmetric,
    diagonally dominant like a relaxed momentum matrix"""
    rng = np.random.default_rng(seed)
    nC, nI, nB = m["nCells"], len(own), len(bfc)
    phi = rng.normal(size=nI) * m["magSf"][:nI]
    D, U, L = np.zeros(nC), np.zeros(nI), np.zeros(nI)
    O.lib().orc_div_upwind(nC, nI, p(own), p(nei), p(f64(phi)), p(D), p(U), p(L))
    lD, lU = np.zeros(nC), np.zeros(nI)
    O.lib().orc_laplacian(nC, nI, p(own), p(nei), p(f64(m["deltaCoeffs"])), p(np.full(nI, 0.05)),
                          p(f64(m["magSf"][:nI])), p(lD), p(lU), None)
    D = D - lD + 0.2 * np.abs(U).max()
    U, L = U - lU, L - lU
    iC = rng.uniform(0.0, 0.5, nB)
    bC = rng.normal(size=nB)
    src = rng.normal(size=nC) * m["V"]
    return D, U, L, src, iC, bC


def dense(nC, own, nei, d, u, l):
    A = np.diag(np.asarray(d, float))
    for f in range(len(own)):
        A[own[f], nei[f]] += u[f]
        A[nei[f], own[f]] += l[f]
    return A


@pytest.mark.skipif(not os.path.exists(REF), reason="reference host library not built (oracle/_ref)")
def test_pbicgstab_bit_exact_with_reference():
    L = ctypes.CDLL(REF)
    L.ref_pbicgstab.restype = ctypes.c_int
    L.ref_pbicgstab.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 6 + [ctypes.c_int] + \
        [ctypes.c_void_p] * 4 + [ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    for seed, (tol, relTol, maxIter, minIter) in ((1, (1e-12, 0.0, 1000, 0)), (2, (1e-3, 0.1, 1000, 0)),
                                                  (3, (1e-30, 0.0, 7, 0)), (4, (1.0, 0.0, 1000, 3))):
        m, own, nei, bfc = channel(9, 7, 5, seed)
        D, U, Lo, src, iC, bC = momentum_system(m, own, nei, bfc, seed)
        nC = m["nCells"]
        x0 = np.random.default_rng(seed + 9).normal(size=nC)
        xo = x0.copy()
        pf = O.lib().orc_pbicgstab(nC, len(own), p(own), p(nei), p(f64(D)), p(f64(U)), p(f64(Lo)), p(f64(src)),
                                   len(bfc), p(bfc), p(f64(iC)), p(f64(bC)), p(xo), tol, relTol, maxIter, minIter)
        xr = x0.copy()
        perf = np.zeros(3)
        assert L.ref_pbicgstab(nC, len(own), p(own), p(nei), p(f64(D)), p(f64(U)), p(f64(Lo)), p(f64(src)),
                               len(bfc), p(bfc), p(f64(iC)), p(f64(bC)), p(xr), tol, relTol, maxIter, minIter,
                               p(perf)) == 0
        assert np.array_equal(xo, xr), seed
        assert pf.initialResidual == perf[0] and pf.finalResidual == perf[1] and pf.nIterations == perf[2], seed


def test_pbicgstab_solves_the_folded_system():
    m, own, nei, bfc = channel(8, 6, 4, 5)
    D, U, Lo, src, iC, bC = momentum_system(m, own, nei, bfc, 5)
    nC = m["nCells"]
    x = np.zeros(nC)
    pf = O.lib().orc_pbicgstab(nC, len(own), p(own), p(nei), p(f64(D)), p(f64(U)), p(f64(Lo)), p(f64(src)),
                               len(bfc), p(bfc), p(f64(iC)), p(f64(bC)), p(x), 1e-14, 0.0, 1000, 0)
    Dc, b = D.copy(), src.copy()
    np.add.at(Dc, bfc, iC)
    np.add.at(b, bfc, bC)
    ref = np.linalg.solve(dense(nC, own, nei, Dc, U, Lo), b)
    assert np.abs(x - ref).max() <= 1e-10 * np.abs(ref).max()
    assert pf.finalResidual < 1e-14


@pytest.mark.skipif(not os.path.exists(REF), reason="reference host library not built (oracle/_ref)")
def test_gamg_pairing_pass_equals_reference_pass():
    """gamg.cu's agglomeratePair walk is the faceAreaPair pass of pair_gamg_agglomeration_cpp.cu: with the
    same (magSf) weights and direction the two give the same map, level after level"""
    L = ctypes.CDLL(REF)
    L.ref_pair_agglomerate.restype = ctypes.c_int
    L.ref_pair_agglomerate.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 5
    for seed in (1, 2):
        m, own, nei, bfc = channel(13, 9, 6, seed)
        nC, lo, up, w = m["nCells"], own, nei, f64(m["magSf"][:len(own)])
        fwd = 1
        D = np.random.default_rng(seed).uniform(-3, -1, nC)
        U = np.random.default_rng(seed + 1).uniform(0.1, 1, len(lo))
        for level in range(6):
            a = np.zeros(nC, np.int32)
            n_o = O.lib().orc_gamg_ref_pair(nC, len(lo), p(i32(lo)), p(i32(up)), p(w), fwd, p(a))
            b = np.zeros(nC, np.int32)
            f = ctypes.c_int(fwd)
            n_r = L.ref_pair_agglomerate(nC, len(lo), p(i32(lo)), p(i32(up)), p(w), ctypes.byref(f), p(b))
            assert n_o == n_r and np.array_equal(a, b), (seed, level)
            if n_o < 4:
                break
            # the next level through the restatement's coarsening
            nCF = O.lib().orc_gamg_ref_coarsen(nC, len(lo), p(i32(lo)), p(i32(up)), p(f64(D)), p(f64(U)), p(f64(U)),
                                               p(w), p(a), n_o, None, None, None, None, None, None)
            cl, cu = np.zeros(nCF, np.int32), np.zeros(nCF, np.int32)
            cD, cUv, cLv, cw = np.zeros(n_o), np.zeros(nCF), np.zeros(nCF), np.zeros(nCF)
            O.lib().orc_gamg_ref_coarsen(nC, len(lo), p(i32(lo)), p(i32(up)), p(f64(D)), p(f64(U)), p(f64(U)), p(w),
                                         p(a), n_o, p(cl), p(cu), p(cD), p(cUv), p(cLv), p(cw))
            nC, lo, up, w, D, U = n_o, cl, cu, cw, cD, cUv
            fwd = 1 - fwd


def test_gamg_coarsening_is_the_galerkin_product():
    """integer-valued asymmetric fine matrix: the coarse level equals R A R^T exactly, its faces are the
    distinct coarse pairs in sorted order, and the face weights sum"""
    m, own, nei, bfc = channel(7, 5, 4, 3, jitter=0.0)
    nC, nI = m["nCells"], len(own)
    rng = np.random.default_rng(3)
    D = rng.integers(-40, -10, nC).astype(float)
    U = rng.integers(1, 9, nI).astype(float)
    Lo = rng.integers(1, 9, nI).astype(float)
    w = rng.integers(1, 5, nI).astype(float)
    a = np.zeros(nC, np.int32)
    nCo = O.lib().orc_gamg_ref_pair(nC, nI, p(own), p(nei), p(f64(m["magSf"][:nI])), 1, p(a))
    nCF = O.lib().orc_gamg_ref_coarsen(nC, nI, p(own), p(nei), p(D), p(U), p(Lo), p(w), p(a), nCo,
                                       None, None, None, None, None, None)
    cl, cu = np.zeros(nCF, np.int32), np.zeros(nCF, np.int32)
    cD, cU, cL, cw = np.zeros(nCo), np.zeros(nCF), np.zeros(nCF), np.zeros(nCF)
    O.lib().orc_gamg_ref_coarsen(nC, nI, p(own), p(nei), p(D), p(U), p(Lo), p(w), p(a), nCo, p(cl), p(cu), p(cD),
                                 p(cU), p(cL), p(cw))
    R = np.zeros((nCo, nC))
    R[a, np.arange(nC)] = 1.0
    Ac = R @ dense(nC, own, nei, D, U, Lo) @ R.T
    assert np.array_equal(dense(nCo, cl, cu, cD, cU, cL), Ac)
    pairs = sorted({(min(a[o], a[n]), max(a[o], a[n])) for o, n in zip(own, nei) if a[o] != a[n]})
    assert [tuple(x) for x in zip(cl, cu)] == pairs
    W = R @ dense(nC, own, nei, np.zeros(nC), w, w) @ R.T
    for f in range(nCF):
        assert cw[f] == W[cl[f], cu[f]]


def test_gamg_gauss_seidel_closed_form():
    m, own, nei, bfc = channel(6, 5, 3, 4)
    D, U, Lo, src, iC, bC = momentum_system(m, own, nei, bfc, 4)
    nC = m["nCells"]
    A = dense(nC, own, nei, D, U, Lo)
    x = np.random.default_rng(8).normal(size=nC)
    y = x.copy()
    O.lib().orc_gamg_ref_gauss_seidel(nC, len(own), p(own), p(nei), p(f64(D)), p(f64(U)), p(f64(Lo)), p(f64(src)),
                                      p(y), 1)
    Lm, Um = np.tril(A), np.triu(A, 1)
    ref = np.linalg.solve(Lm, src - Um @ x)
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()


def test_gamg_dense_and_scale():
    m, own, nei, bfc = channel(4, 3, 3, 6)
    D, U, Lo, src, iC, bC = momentum_system(m, own, nei, bfc, 6)
    nC = m["nCells"]
    A = dense(nC, own, nei, D, U, Lo)
    x = np.zeros(nC)
    O.lib().orc_gamg_ref_dense(nC, len(own), p(own), p(nei), p(f64(D)), p(f64(U)), p(f64(Lo)), p(f64(src)), p(x))
    ref = np.linalg.solve(A, src)
    assert np.abs(x - ref).max() <= 1e-12 * np.abs(ref).max()
    e = np.random.default_rng(2).normal(size=nC)
    got = e.copy()
    O.lib().orc_gamg_ref_scale(nC, len(own), p(own), p(nei), p(f64(D)), p(f64(U)), p(f64(Lo)), p(f64(src)), p(got))
    Ae = A @ e
    sf = (e @ src) / (e @ Ae)
    want = sf * e + (src - sf * Ae) / D
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_gamg_solves_a_pressure_system():
    m, own, nei, bfc = channel(16, 10, 6, 7)
    nC, nI, nB = m["nCells"], len(own), len(bfc)
    gam = np.random.default_rng(7).uniform(0.5, 2.0, nI)
    D, U = np.zeros(nC), np.zeros(nI)
    O.lib().orc_laplacian(nC, nI, p(own), p(nei), p(f64(m["deltaCoeffs"])), p(f64(gam)), p(f64(m["magSf"][:nI])),
                          p(D), p(U), None)
    iC = np.zeros(nB)
    bC = np.zeros(nB)
    out = [k for k, (name, s0, sz) in enumerate(m["patches"]) if name == "outlet"][0]
    _, s0, sz = m["patches"][out]
    sl = slice(s0 - nI, s0 - nI + sz)
    iC[sl] = -m["bDeltaCoeffs"][sl] * m["magSf"][s0:s0 + sz]
    src = np.random.default_rng(8).normal(size=nC) * m["V"]
    x = np.zeros(nC)
    nLev = ctypes.c_int()
    pf = O.lib().orc_gamg_ref(nC, nI, p(own), p(nei), p(D), p(U), p(U), p(f64(src)), nB, p(bfc), p(iC), p(bC),
                              p(f64(m["magSf"][:nI])), p(x), 1e-10, 0.0, 200, 0, 10, 2, 2, ctypes.byref(nLev))
    Dc = D.copy()
    np.add.at(Dc, bfc, iC)
    ref = np.linalg.solve(dense(nC, own, nei, Dc, U, U), src)
    assert nLev.value >= 3
    assert pf.finalResidual < 1e-10 and pf.nIterations < 100
    assert np.abs(x - ref).max() <= 1e-8 * np.abs(ref).max()


def test_grad_u_boundary_linear_field():
    """U = a + B x on a uniform box, B non-symmetric: on a fixedValue face (exact face values) the boundary
    gradient is B; on a zeroGradient face it is the cell gradient with its normal row replaced by zero"""
    m = hexmesh.box_mesh(5, 4, 3, (1.0, 1.0, 1.0), (1.0, 1.0, 1.0), 0.0, seed=0)
    nC, nI = m["nCells"], m["nInternal"]
    bfc = i32(m["owner"][nI:])
    nB = len(bfc)
    B = np.array([[0.3, -1.1, 0.7], [2.0, 0.1, -0.4], [0.9, 1.3, -0.5]])   # B[j, i] = dU_j/dx_i
    a = np.array([1.0, -2.0, 0.5])
    U = [a[j] + m["C"] @ B[j] for j in range(3)]
    Ub = [a[j] + m["Cf"][nI:] @ B[j] for j in range(3)]
    gradU = np.zeros(9 * nC)
    for i in range(3):
        for j in range(3):
            gradU[(i * 3 + j) * nC:(i * 3 + j + 1) * nC] = B[j, i]
    uType = np.full(nB, O.BC_FIXED_VALUE, np.int32)
    uType[::3] = O.BC_ZERO_GRADIENT
    gb = np.zeros(9 * nB)
    O.lib().orc_grad_u_boundary(nC, nI, nB, p(bfc), p(uType), p(f64(m["Sf"])), p(f64(m["magSf"])),
                                p(f64(m["bDeltaCoeffs"])), *[p(f64(u)) for u in U], *[p(f64(u)) for u in Ub],
                                p(gradU), p(gb))
    gb = gb.reshape(9, nB)
    for f in range(nB):
        n = m["Sf"][nI + f] / m["magSf"][nI + f]
        G = B.T.copy()   # G[i, j] = dU_j/dx_i
        if uType[f] == O.BC_ZERO_GRADIENT:
            G = G - np.outer(n, n @ G)
        assert np.allclose(gb[:, f].reshape(3, 3), G, rtol=0, atol=1e-12), f
