"""Pins of the oracle's explicit viscous-stress pieces (linearViscousStress_cpp.cu:36-100, fvc.cu:867-905)
by what the mathematics fixes:

- nu*dev2(T(grad U)) in closed form: an isotropic gradient a*I gives -nu*a*I (dev2's trace is -tr), a
  pure shear d(U_y)/dx = s lands in the (y, x) entry only (the transpose), and the trace of the result
  is -nu*tr(grad U);
- the Gauss divergence of a tensor field: a uniform tensor gives zero in every cell (each cell's face
  areas sum to zero); on a uniform orthogonal box a linear tensor field, with its exact face values on
  the boundary, gives V times the exact divergence d_j = sum_i dT_ij/dx_i -- a non-symmetric gradient,
  so contracting the wrong index fails; empty faces contribute nothing."""
import numpy as np

import oracle_lib as O
from brae_amd import hexmesh


def test_scaled_dev2t_closed_forms():
    n = 4
    nu = np.array([1.0, 2.0, 0.5, 3.0])
    g = np.zeros((9, n))
    g[0] = g[4] = g[8] = [1.0, 2.0, -1.0, 0.25]          # a*I
    sig = O.scaled_dev2t(g, nu)
    a = g[0]
    for q in range(9):
        expect = -nu * a if q in (0, 4, 8) else 0.0
        assert np.allclose(sig[q], expect, rtol=1e-15, atol=0), q
    g = np.zeros((9, n))
    g[0 * 3 + 1] = 3.0                                    # d(U_y)/dx = 3: gradU_xy
    sig = O.scaled_dev2t(g, nu)
    assert np.array_equal(sig[1 * 3 + 0], 3.0 * nu)       # T(gradU)_yx
    assert not np.delete(sig, 3, axis=0).any()
    rng = np.random.default_rng(3)
    g = rng.normal(size=(9, 50))
    nu = rng.uniform(0.1, 2, 50)
    sig = O.scaled_dev2t(g, nu)
    assert np.abs(sig[0] + sig[4] + sig[8] + nu * (g[0] + g[4] + g[8])).max() <= 1e-14 * np.abs(g).max() * 4
    # the off-diagonal entries are the transpose's, scaled
    assert np.allclose(sig[1], nu * g[3], rtol=1e-15, atol=0) and np.allclose(sig[5], nu * g[7], rtol=1e-15, atol=0)


def box(jitter):
    m = hexmesh.box_mesh(6, 5, 4, length=(1.2, 1.0, 0.8), jitter=jitter, seed=2)
    nI = m["nInternal"]
    return m, O.i32(m["owner"][:nI]), O.i32(m["neighbour"]), O.i32(m["owner"][nI:])


def test_uniform_tensor_has_no_divergence():
    m, own, nei, bfc = box(0.12)
    nC, nB = m["nCells"], len(bfc)
    T = np.arange(1.0, 10.0)[:, None]
    d = O.tensor_div(nC, own, nei, m["weights"], m["Sf"], bfc, np.repeat(T, nC, 1), np.repeat(T, nB, 1))
    for j in range(3):
        assert np.abs(d[j]).max() <= 1e-14 * 9 * np.abs(m["Sf"]).max(), j


def test_linear_tensor_divergence_is_exact_on_a_uniform_box():
    m, own, nei, bfc = box(0.0)
    nC, nI = m["nCells"], m["nInternal"]
    rng = np.random.default_rng(5)
    T0, G = rng.normal(size=9), rng.normal(size=(9, 3))    # T_q(x) = T0_q + G_q . x
    Tc = (T0[:, None] + G @ m["C"].T)
    Tb = (T0[:, None] + G @ m["Cf"][nI:].T)
    d = O.tensor_div(nC, own, nei, m["weights"], m["Sf"], bfc, Tc, Tb)
    for j in range(3):
        exact = sum(G[i * 3 + j, i] for i in range(3))        # d_j = sum_i dT_ij/dx_i
        assert np.abs(d[j] - m["V"] * exact).max() <= 1e-12 * np.abs(G).max() * m["V"].max(), j
    # the same field with T's indices swapped in the gradient gives a different answer
    swapped = sum(G[j * 3 + i, i] for i in range(3) for j in [0])
    assert abs(sum(G[i * 3 + 0, i] for i in range(3)) - swapped) > 1e-3
    # empty faces: the boundary's contribution drops out
    empty = np.ones(len(bfc), np.int32)
    dE = O.tensor_div(nC, own, nei, m["weights"], m["Sf"], bfc, Tc, Tb, isEmpty=empty)
    d0 = O.tensor_div(nC, own, nei, m["weights"], m["Sf"], bfc, Tc, np.zeros_like(Tb))
    for j in range(3):
        assert np.array_equal(dE[j], d0[j]), j
