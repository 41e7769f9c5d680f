"""The oracle's Gauss-Seidel sweep and DIC preconditioner (the routines its OpenFOAM GAMG and PCG use),
pinned by closed forms. CPU only.

- One forward Gauss-Seidel sweep is x_new = (D + L)^-1 (b - U x_old): a lower-triangular solve.
- symGaussSeidel's reverse walk keeps the forward walk's lower-neighbour terms:
  x_rev = (D + U)^-1 (b - L x_fwd), an upper-triangular solve (smooth_solver_cpp.cu:141-163).
- DIC on a chain (tridiagonal matrix, no fill-in) is the exact Cholesky solve: w = A^-1 r.
- On a 3-D mesh, DIC's factor satisfies the incomplete-Cholesky identity: (L_D) D^-1 (L_D)^T agrees
  with A on the diagonal and on every face, L_D = D + lower part of A with D = 1/rD.
"""
import numpy as np
from scipy.linalg import solve_triangular

import oracle_lib as O
from brae_amd import hexmesh


def system(shape, seed, jitter=0.1):
    m = hexmesh.box_mesh(*shape, length=(4.0, 1.0, 0.5), grading=(1.0, 20.0, 20.0), jitter=jitter, seed=seed)
    nI, nC = m["nInternal"], m["nCells"]
    own = np.ascontiguousarray(m["owner"][:nI])
    nei = m["neighbour"]
    rng = np.random.default_rng(seed)
    up = -rng.uniform(0.5, 2.0, nI)
    diag = np.zeros(nC)
    np.add.at(diag, own, -up)
    np.add.at(diag, nei, -up)
    diag += rng.uniform(0.1, 1.0, nC)   # strictly diagonally dominant
    return nC, own, nei, diag, up


def dense(nC, own, nei, diag, up):
    A = np.diag(diag)
    A[own, nei] = up
    A[nei, own] = up
    return A


def test_gauss_seidel_sweep_is_triangular_solve():
    nC, own, nei, diag, up = system((7, 5, 4), 1)
    A = dense(nC, own, nei, diag, up)
    rng = np.random.default_rng(2)
    b, x0 = rng.normal(size=nC), rng.normal(size=nC)
    x1 = O.gauss_seidel(nC, own, nei, diag, up, b, x0, 1)
    ref = solve_triangular(np.tril(A), b - np.triu(A, 1) @ x0, lower=True)
    assert np.abs(x1 - ref).max() < 1e-12 * np.abs(ref).max()
    x2 = O.gauss_seidel(nC, own, nei, diag, up, b, x0, 2)
    ref2 = solve_triangular(np.tril(A), b - np.triu(A, 1) @ ref, lower=True)
    assert np.abs(x2 - ref2).max() < 1e-12 * np.abs(ref2).max()


def test_sym_gauss_seidel_sweep_is_forward_then_upper_solve():
    nC, own, nei, diag, up = system((7, 5, 4), 6)
    A = dense(nC, own, nei, diag, up)
    rng = np.random.default_rng(7)
    b, x0 = rng.normal(size=nC), rng.normal(size=nC)
    ref = x0
    for _ in range(2):
        fwd = solve_triangular(np.tril(A), b - np.triu(A, 1) @ ref, lower=True)
        ref = solve_triangular(np.triu(A), b - np.tril(A, -1) @ fwd, lower=False)
        x = O.gauss_seidel(nC, own, nei, diag, up, b, x0, _ + 1, symmetric=True)
        assert np.abs(x - ref).max() < 1e-12 * np.abs(ref).max()
    # a reverse walk restarted from b (dropping the lower terms) is a different sweep
    wrong = solve_triangular(np.triu(A), b, lower=False)
    x1 = O.gauss_seidel(nC, own, nei, diag, up, b, x0, 1, symmetric=True)
    assert np.abs(x1 - wrong).max() > 1e-3 * np.abs(wrong).max()


def test_dic_is_exact_on_a_chain():
    nC, own, nei, diag, up = system((40, 1, 1), 3, jitter=0.0)
    A = dense(nC, own, nei, diag, up)
    r = np.random.default_rng(4).normal(size=nC)
    w = O.dic_precondition(nC, own, nei, diag, up, r)
    ref = np.linalg.solve(A, r)
    assert np.abs(w - ref).max() < 1e-12 * np.abs(ref).max()


def test_dic_factor_matches_the_matrix_on_its_pattern():
    nC, own, nei, diag, up = system((6, 5, 4), 5)
    A = dense(nC, own, nei, diag, up)
    # recover M^-1 column by column, then M = (D+L) D^-1 (D+L)^T must agree with A on diag and faces
    Minv = np.stack([O.dic_precondition(nC, own, nei, diag, up, e) for e in np.eye(nC)], 1)
    M = np.linalg.inv(Minv)
    assert np.abs(np.diag(M) - diag).max() < 1e-9 * np.abs(diag).max()
    assert np.abs(M[own, nei] - up).max() < 1e-9 * np.abs(up).max()
    assert np.abs(M - M.T).max() < 1e-9 * np.abs(M).max()
