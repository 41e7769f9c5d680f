"""Pins of the oracle's fvMatrix::relax and GeometricField::relax (fv_matrix_ops.cuh:158-207,
pEqn_cpp.cu:295-300) by what the mathematics fixes, independently of the formulas the oracle restates:

- the relaxed diagonal, with the removed cmptMin terms added back and times alpha, is the larger of
  |diag + sum cmptMagMax| and the row's off-diagonal magnitude -- the latter from scipy's row sums of the
  assembled asymmetric matrix, so a swapped upper/lower or a dropped face fails;
- relaxation keeps the fixed point: the solution of the unrelaxed system, used as the previous iterate,
  solves the relaxed system too (the source update carries exactly the diagonal's change);
- cmptMagMax/cmptMin of per-component internalCoeffs equal numpy's reductions, and differ where the
  components differ (slip-like faces);
- alpha <= 0 leaves the matrix as it is; relaxField's closed forms."""
import numpy as np
import scipy.sparse.linalg as spla

import oracle_lib as O
from test_oracle_predictor import dense, momentum_system


def off_row_sums(nC, own, nei, upper, lower):
    A = dense(nC, own, nei, np.zeros(nC), upper, lower)
    return np.asarray(abs(A).sum(axis=1)).ravel()


def per_cell(nC, bfc, v):
    out = np.zeros(nC)
    np.add.at(out, bfc, v)
    return out


def test_relaxed_diagonal_is_the_dominance_bound():
    m, own, nei, bfc, diag, upper, lower, ic, bc, src = momentum_system()
    nC = m["nCells"]
    assert np.abs(upper - lower).max() > 1e-3
    ic[1] = ic[1] * np.where(np.arange(len(bfc)) % 3 == 0, -1.5, 1.0)   # components that differ, some negative
    mag, mn = O.cmpt_mag_max_min(*ic)
    # make some rows lose dominance so both branches of the max occur
    d0 = diag.copy()
    d0[::4] *= 0.05
    off = off_row_sums(nC, own, nei, upper, lower)
    for alpha in (0.7, 0.3, 1.0):
        d, delta, _ = O.relax_matrix(nC, own, nei, upper, lower, bfc, mag, mn, alpha, d0)
        back = (d + per_cell(nC, bfc, mn)) * alpha
        bound = np.maximum(np.abs(d0 + per_cell(nC, bfc, mag)), off)
        assert np.abs(back - bound).max() <= 1e-13 * bound.max(), alpha
        assert (off > np.abs(d0 + per_cell(nC, bfc, mag))).sum() > 10      # the off-diagonal branch is taken
        assert np.array_equal(delta, d - d0)


def test_relaxation_keeps_the_fixed_point():
    m, own, nei, bfc, diag, upper, lower, ic, bc, src = momentum_system()
    nC = m["nCells"]
    mag, mn = O.cmpt_mag_max_min(*ic)
    psiStar = []
    for k in range(3):
        dk = diag + per_cell(nC, bfc, ic[k])
        psiStar.append(spla.spsolve(dense(nC, own, nei, dk, upper, lower).tocsc(), src[k] + per_cell(nC, bfc, bc[k])))
    d, delta, srcR = O.relax_matrix(nC, own, nei, upper, lower, bfc, mag, mn, 0.7, diag, psi=psiStar, source=src)
    assert np.abs(delta).max() > 1e-3                                    # the diagonal did change
    for k in range(3):
        dk = d + per_cell(nC, bfc, ic[k])
        x = spla.spsolve(dense(nC, own, nei, dk, upper, lower).tocsc(), srcR[k] + per_cell(nC, bfc, bc[k]))
        assert np.abs(x - psiStar[k]).max() <= 1e-10 * np.abs(psiStar[k]).max(), k


def test_component_reductions_and_no_op():
    rng = np.random.default_rng(8)
    x, y, z = rng.normal(size=(3, 300))
    mag, mn = O.cmpt_mag_max_min(x, y, z)
    S = np.stack([x, y, z])
    assert np.array_equal(mag, np.abs(S).max(axis=0)) and np.array_equal(mn, S.min(axis=0))
    assert (mag != np.abs(mn)).any()
    # equal components: the added and the removed terms are the same number for ic >= 0
    mag1, mn1 = O.cmpt_mag_max_min(np.abs(x), np.abs(x), np.abs(x))
    assert np.array_equal(mag1, mn1)
    m, own, nei, bfc, diag, upper, lower, ic, bc, src = momentum_system(5, 4, 3, 1)
    nC = m["nCells"]
    for alpha in (0.0, -0.5):
        d, delta, srcR = O.relax_matrix(nC, own, nei, upper, lower, bfc, *O.cmpt_mag_max_min(*ic), alpha, diag,
                                        psi=[np.ones(nC)], source=[src[0]])
        assert np.array_equal(d, diag) and not delta.any() and np.array_equal(srcR[0], src[0])


def test_coupled_interface_magnitude_joins_the_bound():
    m, own, nei, bfc, diag, upper, lower, ic, bc, src = momentum_system(5, 4, 3, 2)
    nC = m["nCells"]
    mag, mn = O.cmpt_mag_max_min(*ic)
    cyc = np.zeros(nC)
    cyc[7] = 1e3
    d, _, _ = O.relax_matrix(nC, own, nei, upper, lower, bfc, mag, mn, 0.5, diag, cycSumOff=cyc)
    off = off_row_sums(nC, own, nei, upper, lower)
    assert abs((d[7] + per_cell(nC, bfc, mn)[7]) * 0.5 - (off[7] + 1e3)) <= 1e-12 * 1e3
    d2, _, _ = O.relax_matrix(nC, own, nei, upper, lower, bfc, mag, mn, 0.5, diag)
    assert np.array_equal(np.delete(d, 7), np.delete(d2, 7))


def test_relax_field_closed_forms():
    p, prev = np.array([1.0, -2.0, 4.0]), np.array([3.0, 2.0, 0.0])
    assert O.relax_field(p, prev, 0.5).tolist() == [2.0, 0.0, 2.0]
    assert O.relax_field(p, prev, 0.25).tolist() == [2.5, 1.0, 1.0]
    for a in (1.0, 1.5, 0.0, -1.0):
        assert np.array_equal(O.relax_field(p, prev, a), p)
