"""The oracle's non-orthogonal correction pinned by what the mathematics fixes (the reference holds no
dump with a non-zero correction source). CPU only.

- Gauss linear gradient of a linear field is exact on an orthogonal graded box (the face value of a
  linear field is exact where cell centres line up with the face centre), with fixedValue patches.
- Divergence theorem on the jittered mesh: sum_c V_c grad_c = sum_b Sf_b p_b (internal faces cancel).
- Corrected face flux of a linear field: gamma*|Sf|*ndc*(p_N - p_P) + ffc = gamma*Sf.a exactly when
  the cell gradient is a (corrVec = n - delta*ndc), on every face of the jittered mesh.
- The source loses exactly the divergence of ffc: it sums to zero over the mesh.
"""
import numpy as np

import oracle_lib as O
from brae_amd import hexmesh

A_VEC = np.array([0.7, -1.3, 2.1])


def mesh(jitter, seed=3, shape=(9, 7, 5)):
    m = hexmesh.box_mesh(*shape, length=(4.0, 1.0, 0.5), grading=(1.0, 20.0, 20.0), jitter=jitter, seed=seed)
    nI = m["nInternal"]
    return m, np.ascontiguousarray(m["owner"][:nI]), m["neighbour"], np.ascontiguousarray(m["owner"][nI:])


def test_grad_of_linear_field_exact_on_graded_box():
    m, own, nei, bfc = mesh(0.0)
    nI, nB = len(own), len(bfc)
    psi = m["C"] @ A_VEC + 0.25
    types = np.full(nB, O.BC_FIXED_VALUE, np.int32)
    vals = m["Cf"][nI:] @ A_VEC + 0.25
    g = O.grad_gauss(m["nCells"], own, nei, bfc, types, vals, m["Sf"], m["weights"], m["V"], psi)
    assert np.abs(g - A_VEC).max() < 1e-9 * np.abs(A_VEC).max()


def test_grad_zero_gradient_and_empty_patches():
    """zeroGradient takes the cell value; an empty patch adds nothing (a uniform field's gradient is then
    the empty faces' missing Sf sum, which is what OpenFOAM gives on an extruded 2-D mesh)"""
    m, own, nei, bfc = mesh(0.1)
    nB = len(bfc)
    psi = np.full(m["nCells"], 3.0)
    g = O.grad_gauss(m["nCells"], own, nei, bfc, np.zeros(nB, np.int32), np.zeros(nB), m["Sf"], m["weights"],
                     m["V"], psi)
    assert np.abs(g).max() < 1e-9
    types = np.zeros(nB, np.int32)
    types[hexmesh.patch_slice(m, "front")] = O.BC_EMPTY
    g2 = O.grad_gauss(m["nCells"], own, nei, bfc, types, np.zeros(nB), m["Sf"], m["weights"], m["V"], psi)
    nI = len(own)
    missing = np.zeros((m["nCells"], 3))
    s = hexmesh.patch_slice(m, "front")
    np.add.at(missing, bfc[s], -3.0 * m["Sf"][nI:][s])
    assert np.abs(g2 - missing / m["V"][:, None]).max() < 1e-9 * np.abs(missing / m["V"][:, None]).max()


def test_grad_divergence_theorem_on_jittered_mesh():
    m, own, nei, bfc = mesh(0.1)
    nI, nB = len(own), len(bfc)
    rng = np.random.default_rng(1)
    psi = rng.normal(size=m["nCells"])
    types = rng.integers(0, 2, nB).astype(np.int32)
    vals = rng.normal(size=nB)
    g = O.grad_gauss(m["nCells"], own, nei, bfc, types, vals, m["Sf"], m["weights"], m["V"], psi)
    pb = np.where(types == O.BC_FIXED_VALUE, vals, psi[bfc])
    lhs = (g * m["V"][:, None]).sum(0)
    rhs = (m["Sf"][nI:] * pb[:, None]).sum(0)
    assert np.abs(lhs - rhs).max() < 1e-10 * np.abs(g * m["V"][:, None]).sum()


def test_corrected_flux_of_linear_field_is_exact():
    m, own, nei, _ = mesh(0.1)
    nI = len(own)
    rng = np.random.default_rng(2)
    gamma = rng.uniform(0.5, 2.0, nI)
    psi = m["C"] @ A_VEC
    grad = np.tile(A_VEC, (m["nCells"], 1))
    ffc, _ = O.laplacian_corr(m["nCells"], own, nei, grad, m["weights"], m["corrVec"], gamma, m["magSf"][:nI])
    assert np.abs(ffc).max() > 1e-3 * np.abs(gamma * m["magSf"][:nI]).max()   # the mesh is non-orthogonal
    implicit = gamma * m["magSf"][:nI] * m["nonOrthDeltaCoeffs"] * (psi[nei] - psi[own])
    exact = gamma * (m["Sf"][:nI] @ A_VEC)
    assert np.abs(implicit + ffc - exact).max() < 1e-10 * np.abs(exact).max()


def test_correction_vanishes_on_orthogonal_mesh_and_source_is_its_divergence():
    m0, own0, nei0, _ = mesh(0.0)
    nI0 = len(own0)
    g0 = np.random.default_rng(3).normal(size=(m0["nCells"], 3))
    ffc0, _ = O.laplacian_corr(m0["nCells"], own0, nei0, g0, m0["weights"], m0["corrVec"], np.ones(nI0),
                               m0["magSf"][:nI0])
    assert np.abs(ffc0).max() < 1e-12
    m, own, nei, _ = mesh(0.1)
    nI = len(own)
    rng = np.random.default_rng(4)
    g = rng.normal(size=(m["nCells"], 3))
    src0 = rng.normal(size=m["nCells"])
    ffc, src = O.laplacian_corr(m["nCells"], own, nei, g, m["weights"], m["corrVec"], rng.uniform(0.5, 2, nI),
                                m["magSf"][:nI], source=src0)
    div = np.zeros(m["nCells"])
    np.add.at(div, own, ffc)
    np.add.at(div, nei, -ffc)
    assert np.abs((src0 - src) - div).max() < 1e-12 * np.abs(ffc).max()
    assert abs((src0 - src).sum()) < 1e-10 * np.abs(ffc).sum()


def test_flux_correct_adds_the_correction():
    m, own, nei, _ = mesh(0.1)
    nI = len(own)
    rng = np.random.default_rng(5)
    up, psi, ffc, phiH = rng.normal(size=nI), rng.normal(size=m["nCells"]), rng.normal(size=nI), rng.normal(size=nI)
    phi = O.flux_correct(own, nei, up, None, psi, ffc, phiH)
    phi0 = O.flux_correct(own, nei, up, None, psi, None, phiH)
    assert np.allclose(phi0 - phi, ffc, rtol=0, atol=1e-12)
    assert np.allclose(phi0, phiH - (up * psi[nei] - up * psi[own]), rtol=0, atol=1e-12)


def test_limited_correction_caps_against_the_orthogonal_part():
    """`limited k corrected`: k outside (0, 1) is the unlimited correction to the bit; inside, a face keeps
    its full correction (to the bit) where k|orth| >= (1-k)|corr|, and elsewhere its flux magnitude is
    exactly gamma|Sf| k|orth| |corr| / ((1-k)|corr| + 1e-15) <= gamma|Sf| k/(1-k) |orth|; the source
    loses exactly the divergence of the limited flux."""
    m, own, nei, bfc = mesh(0.3, seed=5)
    nC, nI = m["nCells"], len(own)
    rng = np.random.default_rng(6)
    vf = rng.normal(size=nC)
    grad = rng.normal(size=(nC, 3)) * 20.0   # a rough gradient: large corrections on the jittered faces
    gam = rng.uniform(0.5, 2.0, nI)
    args = (nC, own, nei, grad, m["weights"], m["corrVec"], gam, m["magSf"][:nI])
    full, _ = O.laplacian_corr(*args)
    for k in (0.0, 1.0, 1.5, -0.2):
        ffc, _ = O.laplacian_corr_limited(*args, m["nonOrthDeltaCoeffs"], vf, k)
        assert np.array_equal(ffc, full), k
    for k in (0.333, 0.5, 0.9):
        ffc, src = O.laplacian_corr_limited(*args, m["nonOrthDeltaCoeffs"], vf, k, source=np.zeros(nC))
        orth = m["nonOrthDeltaCoeffs"] * (vf[nei] - vf[own])
        gm = gam * m["magSf"][:nI]
        corr = np.where(gm != 0, full / gm, 0.0)
        keep = k * np.abs(orth) >= (1 - k) * np.abs(corr) + 1e-15
        assert keep.any() and (~keep).any(), k
        assert np.array_equal(ffc[keep], full[keep])
        capped = np.abs(gm[~keep]) * k * np.abs(orth[~keep]) * np.abs(corr[~keep]) / (
            (1 - k) * np.abs(corr[~keep]) + 1e-15)
        assert np.allclose(np.abs(ffc[~keep]), capped, rtol=1e-12, atol=0)
        assert (np.abs(ffc) <= np.abs(gm) * k / (1 - k) * np.abs(orth) * (1 + 1e-12) + np.abs(full) * keep).all()
        div = np.zeros(nC)
        np.add.at(div, own, ffc)
        np.add.at(div, nei, -ffc)
        assert np.allclose(src, -div, rtol=0, atol=1e-12 * np.abs(ffc).max())
