"""The oracle pinned against the reference's own golden data (OpenFOAM v2412 dumps and the
20-cell AMG trace), at the reference tests' own tolerances. CPU only."""
import os

import numpy as np
import pytest

import oracle_lib as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def pitz():
    d = np.load(os.path.join(GOLD, "pitzdaily.npz"))
    return d, O.Mesh.from_npz(d)


def rel(a, b):
    m = np.abs(b).max()
    return np.abs(a - b).max() / (m if m > 0 else 1.0)


def t_boundary(d, mesh):
    """T's patch fields on pitzDaily: inlet fixedValue 1, outlet fixedValue 3, walls zeroGradient,
    frontAndBack empty (validation/matrixDump/282/T)."""
    types = np.zeros(mesh.nBFaces, np.int32)
    vals = d["T_values"].copy()
    for name, t in zip(mesh.patchNames, d["T_types"]):
        s = mesh.patch_slice(name)
        types[s] = {"fixedValue": O.BC_FIXED_VALUE, "zeroGradient": O.BC_ZERO_GRADIENT, "empty": O.BC_EMPTY}[str(t)]
    return types, vals


def t_laplacian(d, mesh):
    nI = mesh.nInternal
    # tests/test_fvm_laplacian.cu: fvm::laplacian(T, 1.0) (orthogonal deltaCoeffs)
    diag, upper, lower = O.laplacian(mesh, mesh.deltaCoeffs, np.ones(nI))
    types, vals = t_boundary(d, mesh)
    iC, bC = O.laplacian_bcoeffs(mesh, types, vals, np.ones(mesh.nBFaces))
    return diag, upper, lower, types, iC, bC


def test_laplacian_matches_openfoam_dump(pitz):
    """test_fvm_laplacian.cu:20-31: diag/upper/lower and per-patch iC/bC at rel <= 1e-12."""
    d, mesh = pitz
    diag, upper, lower, types, iC, bC = t_laplacian(d, mesh)
    assert rel(diag, d["lap_diag"]) <= 1e-12
    assert rel(upper, d["lap_upper"]) <= 1e-12
    assert rel(lower, d["lap_lower"]) <= 1e-12
    for name in mesh.patchNames:
        ref_i, ref_b = d["lap_iC_" + name], d["lap_bC_" + name]
        if ref_i.size == 0:
            continue
        s = mesh.patch_slice(name)
        assert rel(iC[s], ref_i) <= 1e-12, name
        assert rel(bC[s], ref_b) <= 1e-12, name


def test_div_upwind_matches_openfoam_dump(pitz):
    """test_fvm_div.cu: fvm::div(phi, T) Gauss upwind against matrix_div.dat."""
    d, mesh = pitz
    nC, nI = mesh.nCells, mesh.nInternal
    diag, upper, lower = np.zeros(nC), np.zeros(nI), np.zeros(nI)
    phi = O.f64(d["phi_internal"])
    O.lib().orc_div_upwind(nC, nI, O.p(mesh.own()), O.p(mesh.neighbour), O.p(phi), O.p(diag), O.p(upper), O.p(lower))
    assert rel(diag, d["div_diag"]) <= 1e-12
    assert rel(upper, d["div_upper"]) <= 1e-12
    assert rel(lower, d["div_lower"]) <= 1e-12
    types, vals = t_boundary(d, mesh)
    iC, bC = np.zeros(mesh.nBFaces), np.zeros(mesh.nBFaces)
    O.lib().orc_div_bcoeffs(mesh.nBFaces, O.p(types), O.p(O.f64(vals)), O.p(O.f64(d["phi_boundary"])), O.p(iC), O.p(bC))
    for name in mesh.patchNames:
        ref_i = d["div_iC_" + name]
        if ref_i.size == 0:
            continue
        s = mesh.patch_slice(name)
        assert rel(iC[s], ref_i) <= 1e-12, name
        assert rel(bC[s], d["div_bC_" + name]) <= 1e-12, name


def test_div_linear_pinned_by_upwind_and_divergence(pitz):
    """Gauss linear convection (deviceDivCentralCoeffs): with the weights set to pos0(phi) it is the upwind
    matrix that matches OpenFOAM's dump above, bit for bit; with the mesh's geometric weights (any
    weights) the matrix applied to a constant field is the internal-face divergence of phi; and at the
    weight's end points a face takes the neighbour's (w = 0) or the owner's (w = 1) value alone."""
    d, mesh = pitz
    nC, nI = mesh.nCells, mesh.nInternal
    own, nei = mesh.own(), mesh.neighbour
    phi = O.f64(d["phi_internal"])

    def lin(w):
        diag, upper, lower = np.zeros(nC), np.zeros(nI), np.zeros(nI)
        O.lib().orc_div_linear(nC, nI, O.p(own), O.p(nei), O.p(phi), O.p(O.f64(w)), O.p(diag), O.p(upper), O.p(lower))
        return diag, upper, lower

    ud, uu, ul = lin((phi >= 0).astype(np.float64))
    assert rel(ud, d["div_diag"]) <= 1e-12 and rel(uu, d["div_upper"]) <= 1e-12 and rel(ul, d["div_lower"]) <= 1e-12
    ref = [np.zeros(nC), np.zeros(nI), np.zeros(nI)]
    O.lib().orc_div_upwind(nC, nI, O.p(own), O.p(nei), O.p(phi), *[O.p(a) for a in ref])
    assert all(np.array_equal(a, b) for a, b in zip((ud, uu, ul), ref))
    diag, upper, lower = lin(mesh.weights)
    one = O.amul(nC, own, nei, diag, upper, lower, np.ones(nC))
    div = np.zeros(nC)
    np.add.at(div, own, phi)
    np.add.at(div, nei, -phi)
    assert np.abs(one - div).max() <= 1e-12 * np.abs(phi).max()
    # the end points of the weight: w = 0 takes the neighbour's value (lower 0, upper phi), w = 1 the owner's
    d0, u0, l0 = lin(np.zeros(nI))
    d1, u1, l1 = lin(np.ones(nI))
    assert np.array_equal(l0, np.zeros(nI)) and np.array_equal(u0, phi)
    assert np.array_equal(l1, -phi) and np.array_equal(u1, np.zeros(nI))


def test_amul_matches_openfoam_dump(pitz):
    """test_spmv.cu:42-47: A.psi against amul.dat at rel <= 1e-12 (A = laplacian(1,T), unfolded)."""
    d, mesh = pitz
    diag, upper, lower, *_ = t_laplacian(d, mesh)
    Apsi = O.amul(mesh.nCells, mesh.own(), mesh.neighbour, diag, upper, lower, d["amul_psi"])
    assert rel(Apsi, d["amul_Apsi"]) <= 1e-12


def test_amul_equals_dense_product(pitz):
    """A.psi by the face loop equals the dense matrix-vector product of the same LDU matrix."""
    rng = np.random.default_rng(3)
    n, own, nei = 40, [], []
    for c in range(n - 1):
        own.append(c); nei.append(c + 1)
        if c + 7 < n:
            own.append(c); nei.append(c + 7)
    own, nei = np.array(own, np.int32), np.array(nei, np.int32)
    diag, upper, lower = rng.normal(size=n), rng.normal(size=own.size), rng.normal(size=own.size)
    A = np.diag(diag)
    A[own, nei] += upper
    A[nei, own] += lower
    x = rng.normal(size=n)
    np.testing.assert_allclose(O.amul(n, own, nei, diag, upper, lower, x), A @ x, rtol=0, atol=1e-12)


def solve_system(d, mesh):
    """-laplacian(1,T) folded, as dumpSolve builds it (test_pcg.cu:31-38)."""
    diag, upper, lower, types, iC, bC = t_laplacian(d, mesh)
    diagC, b = O.fold(mesh, -diag, np.zeros(mesh.nCells), types, -iC, -bC)
    return diagC, -upper, b


def test_pcg_dic_matches_openfoam_iterations(pitz):
    """test_pcg.cu:59-61: 224 iterations exactly, initial residual 1e-9, solution 1e-6."""
    d, mesh = pitz
    diagC, upper, b = solve_system(d, mesh)
    psi, perf = O.pcg_dic(mesh.nCells, mesh.own(), mesh.neighbour, diagC, upper, b, np.zeros(mesh.nCells), 1e-8)
    assert perf.nIterations == int(d["pcg_iters"]) == 224
    assert abs(perf.initialResidual - float(d["pcg_init"])) / float(d["pcg_init"]) < 1e-9
    assert rel(psi, d["pcg_T"]) < 1e-6


def test_gamg_matches_openfoam_iterations(pitz):
    """OpenFOAM GAMG (faceAreaPair, GaussSeidel, V(0,2,2), scaled, PCG-DIC coarsest) on the
    matrixDumpGAMG case: OpenFOAM took 101 V-cycles to 1e-8; the restatement must take the same
    number and agree on the solution to 1e-6 (test_gamg.cu:58-59)."""
    d, mesh = pitz
    diagC, upper, b = solve_system(d, mesh)
    w = O.faceareapair_weights(mesh.Sf[:mesh.nInternal], mesh.magSf[:mesh.nInternal])
    ag = O.of_agglomeration(mesh.nCells, mesh.own(), mesh.neighbour, w, nCellsInCoarsestLevel=10)
    psi, perf = O.gamg_solve(ag, diagC, upper, b, np.zeros(mesh.nCells), tolerance=1e-8, relTol=0.0,
                             smoother="GaussSeidel")
    assert perf.finalResidual < 1e-8
    assert rel(psi, d["gamg_T"]) < 1e-6
    assert perf.nIterations == int(d["gamg_iters"]) == 101


def test_pcg_gamg_converges_to_openfoam_solution(pitz):
    d, mesh = pitz
    diagC, upper, b = solve_system(d, mesh)
    w = O.faceareapair_weights(mesh.Sf[:mesh.nInternal], mesh.magSf[:mesh.nInternal])
    ag = O.of_agglomeration(mesh.nCells, mesh.own(), mesh.neighbour, w, nCellsInCoarsestLevel=10)
    psi, perf = O.pcg_gamg_solve(ag, diagC, upper, b, np.zeros(mesh.nCells), 1e-8, smoother="GaussSeidel")
    assert perf.finalResidual < 1e-8
    assert perf.nIterations < 40
    assert rel(psi, d["pcg_T"]) < 1e-6


# ---- 20-cell AMG trace (demo/amgpcg/case20) ----

@pytest.fixture(scope="module")
def c20():
    d = np.load(os.path.join(GOLD, "case20.npz"))
    return d, O.Mesh.from_npz(d)


def test_case20_matrix_is_mesh_laplacian(c20):
    d, mesh = c20
    assert np.array_equal(d["m_owner"], mesh.own()) and np.array_equal(d["m_neighbour"], mesh.neighbour)


def test_of_faceareapair_hierarchy_bit_exact(c20):
    """of_hierarchy.csv is OpenFOAM's own faceAreaPair agglomeration (nCellsInCoarsestLevel 4)."""
    d, mesh = c20
    w = O.faceareapair_weights(mesh.Sf[:mesh.nInternal], mesh.magSf[:mesh.nInternal])
    ag = O.of_agglomeration(mesh.nCells, mesh.own(), mesh.neighbour, w, nCellsInCoarsestLevel=4)
    ref = d["of_hierarchy"]
    assert ag.nLevels == ref[:, 0].max() + 1
    for i, L in enumerate(ag.levels):
        rows = ref[ref[:, 0] == i]
        assert np.array_equal(L["map"][rows[:, 1]], rows[:, 2]), i


def test_brae_hierarchy_and_galerkin_bit_exact(c20):
    """brae_hierarchy.csv / brae_levels.csv are brae's buildAMG + amgGalerkin with
    BRAE_AMG_TARGET=4 and |Sf| face weights (trace_brae.cu:404-446)."""
    d, mesh = c20
    ag = O.brae_agglomeration(mesh.nCells, mesh.own(), mesh.neighbour, mesh.magSf[:mesh.nInternal], target=4)
    ref = d["brae_hierarchy"]
    assert ag.nLevels == ref[:, 0].max() + 1
    for i, L in enumerate(ag.levels):
        rows = ref[ref[:, 0] == i]
        assert np.array_equal(L["map"][rows[:, 1]], rows[:, 2]), i
    lv = d["brae_levels"]
    coarse = ag.galerkin(d["m_diag"], d["m_upper"], order="brae")
    for g, (cd, cu) in enumerate(coarse, start=1):
        rows = lv[lv[:, 0] == g]
        dr = rows[rows[:, 1] == 0]
        ur = rows[rows[:, 1] == 1]
        assert np.array_equal(cd[dr[:, 2].astype(int)], dr[:, 5]), g
        L = ag.levels[g - 1]
        fidx = ur[:, 2].astype(int)
        assert np.array_equal(L["cOwn"][fidx], ur[:, 3].astype(np.int32)), g
        assert np.array_equal(L["cNei"][fidx], ur[:, 4].astype(np.int32)), g
        assert np.array_equal(cu[fidx], ur[:, 5]), g


def test_brae_amgpcg_reproduces_device_trace(c20):
    """brae_cycles.csv: psi after each of 30 iterations of the real deviceAMGPCG (FP32 V(1,1)
    Jacobi cycle, direct coarsest, normFactor sum|b|+1e-20). The FP32 cycle agrees with the dump to
    FP32 precision; the FP64 cycle to the FP32 rounding the dump carries."""
    d, mesh = c20
    ag = O.brae_agglomeration(mesh.nCells, mesh.own(), mesh.neighbour, mesh.magSf[:mesh.nInternal], target=4)
    b = d["rhs"]
    nf = np.abs(b).sum() + 1e-20
    cyc = d["brae_cycles"]
    K = int(cyc[:, 0].max())
    _, perf, tr = O.brae_amgpcg(ag, d["m_diag"], d["m_upper"], b, np.zeros(mesh.nCells), 0.0, maxIter=K,
                                single=True, normFactor=nf, trace=True)
    assert perf.nIterations == K
    worst = 0.0
    for k in range(1, 8):
        ref = cyc[cyc[:, 0] == k][:, 2]
        worst = max(worst, np.abs(tr[k] - ref).max())
    assert worst < 2e-6
    # converged state: the solution to machine precision
    np.testing.assert_allclose(tr[K], cyc[cyc[:, 0] == K][:, 2], rtol=0, atol=1e-12)
