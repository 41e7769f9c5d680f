"""The oracle's k-epsilon model (oracle/kepsilon.c), pinned against OpenFOAM's own k-epsilon dumps of the
reference's validation cases and against what the mathematics fixes.

- kEqn and epsilonEqn as assembled matrices, on pitzDaily from OpenFOAM's step-282 fields: diag, upper,
  lower, source and every patch's coefficients against OpenFOAM's keqn.dat/epseqn.dat at the reference's
  1e-11 (tests/test_kepsilon_eqn.cu). A dropped Sp term, a wrong sigma or a swapped diffusivity fails it.
- One correct() from step 282 against OpenFOAM's kEpsilon::correct() (correct.dat): k, epsilon and nut at the
  reference's 6e-3 (the restatement reaches 1e-7), and the near-wall epsilon the wall function fixes at 1e-9 (tests/test_kepsilon_correct.cu).
  That pins the wall functions, the near-wall distance, the production override and setValues.
- Closed forms: yPlusLam's fixed point, the nutkWallFunction's two branches, GbyNu of pure shear (s^2), pure
  extension (4a^2) and a random tensor (t:(t + t^T) - (2/3) tr(t)^2), the near-wall distance on an orthogonal
  graded box (the centre's distance to its wall plane; the smaller of the two at a corner).
"""
import os

import numpy as np
import pytest

import oracle_lib as O
from brae_amd import hexmesh

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

TYPES = {"fixedValue": O.BC_FIXED_VALUE, "noSlip": O.BC_FIXED_VALUE, "zeroGradient": O.BC_ZERO_GRADIENT,
         "kqRWallFunction": O.BC_ZERO_GRADIENT, "epsilonWallFunction": O.BC_ZERO_GRADIENT,
         "calculated": O.BC_ZERO_GRADIENT, "nutkWallFunction": O.BC_ZERO_GRADIENT, "empty": O.BC_EMPTY}


def rel(a, b):
    m = np.abs(b).max()
    return np.abs(a - b).max() / (m if m > 0 else 1.0)


def pitz(tag):
    d = np.load(os.path.join(GOLD, "kepsilon.npz"))
    mesh = O.Mesh(d["mesh_points"], d["mesh_faceOffsets"], d["mesh_faceVerts"], d["mesh_owner"],
                  d["mesh_neighbour"], d["mesh_patchNames"], d["mesh_patchStart"], d["mesh_patchSize"],
                  d["mesh_patchTypes"])
    sizes = mesh.patchSize
    codes = lambda f: O.i32([TYPES[str(t)] for t in np.repeat(d[tag + "_" + f + "_types"], sizes)])
    bfc = mesh.bFaceCell
    uT = codes("U")
    uV = d[tag + "_U_values"]
    U = [O.f64(d[tag + "_U0"])[:, k].copy() for k in range(3)]
    Ub = [O.evaluate_boundary(bfc, uT, uV[:, k], U[k]) for k in range(3)]
    kT, eT = codes("k"), codes("epsilon")
    kV, eV = d[tag + "_k_values"], d[tag + "_epsilon_values"]
    k, eps, nut = (O.f64(d[tag + "_" + f + "0"]).copy() for f in ("k", "epsilon", "nut"))
    kb = O.evaluate_boundary(bfc, kT, kV, k)
    epsb = O.evaluate_boundary(bfc, eT, eV, eps)
    nutb = O.f64(d[tag + "_nut_values"]).copy()
    isWall = O.i32(np.repeat([t == "wall" for t in mesh.patchTypes], sizes))
    return d, mesh, U, Ub, uT, kT, kV, eT, eV, k, kb, eps, epsb, nut, nutb, isWall


def grad_u(mesh, U, Ub, uT):
    gT = np.where(uT == O.BC_EMPTY, O.BC_EMPTY, O.BC_FIXED_VALUE).astype(np.int32)
    n = mesh.nCells
    gradU = np.zeros((9, n))
    for j in range(3):
        g = O.grad_gauss(n, mesh.own(), mesh.neighbour, mesh.bFaceCell, gT, Ub[j], mesh.Sf, mesh.weights, mesh.V,
                         U[j]).reshape(n, 3)
        for i in range(3):
            gradU[i * 3 + j] = g[:, i]
    return gradU


def test_kepsilon_equations_match_openfoam():
    d, mesh, U, Ub, uT, kT, kV, eT, eV, k, kb, eps, epsb, nut, nutb, isWall = pitz("eqn")
    y = np.zeros(mesh.nFaces - mesh.nInternal)
    cs = O.kepsilon_case(mesh, isWall, y, kT, kV, eT, eV, nu=1e-5)
    gbn = O.gbynu(grad_u(mesh, U, Ub, uT))
    G = nut * gbn
    phi, phiB = d["eqn_phi0_internal"], d["eqn_phi0_boundary"]
    kq, eq = O.kepsilon_eqns(cs, k, eps, nut, None, G, gbn, phi, phiB)
    for name, (D, Up, Lo, S, iC, bC) in (("k", kq), ("eps", eq)):
        worst = max(rel(D, d[name + "eqn_diag"]), rel(Up, d[name + "eqn_upper"]), rel(Lo, d[name + "eqn_lower"]),
                    rel(S, d[name + "eqn_source"]))
        for pn, st, sz in zip(mesh.patchNames, mesh.patchStart, mesh.patchSize):
            ref_i, ref_b = d[name + "eqn_iC_" + pn], d[name + "eqn_bC_" + pn]
            if ref_i.size == 0:
                continue
            s = slice(st - mesh.nInternal, st - mesh.nInternal + sz)
            worst = max(worst, rel(iC[s], ref_i), rel(bC[s], ref_b))
        assert worst <= 1e-11, (name, worst)


def test_kepsilon_equations_fail_a_planted_error():
    """the pin above is sensitive: the epsilon equation with sigmaEps = sigmak (1.0) misses OpenFOAM's"""
    d, mesh, U, Ub, uT, kT, kV, eT, eV, k, kb, eps, epsb, nut, nutb, isWall = pitz("eqn")
    cs = O.kepsilon_case(mesh, isWall, np.zeros(mesh.nFaces - mesh.nInternal), kT, kV, eT, eV, nu=1e-5,
                         co=O.kepsilon_coeffs(sigmaEps=1.0))
    gbn = O.gbynu(grad_u(mesh, U, Ub, uT))
    _, eq = O.kepsilon_eqns(cs, k, eps, nut, None, nut * gbn, gbn, d["eqn_phi0_internal"], d["eqn_phi0_boundary"])
    assert rel(eq[1], d["epseqn_upper"]) > 1e-6


def test_kepsilon_correct_matches_openfoam():
    d, mesh, U, Ub, uT, kT, kV, eT, eV, k, kb, eps, epsb, nut, nutb, isWall = pitz("corr")
    y = O.near_wall_dist(mesh, isWall)
    cs = O.kepsilon_case(mesh, isWall, y, kT, kV, eT, eV, nu=1e-5, relaxEps=1.0, relaxK=1.0, tol=1e-10,
                         relTol=0.0, maxIter=2000)
    r = O.kepsilon_correct(cs, U, Ub, d["corr_phi0_internal"], d["corr_phi0_boundary"], k, kb, eps, epsb, nut, nutb)
    assert r.nWallCells > 0 and r.epsIters > 0 and r.kIters > 0
    # the reference's gate is 6e-3; the restatement lands at 1e-7 (k), 7e-14 (epsilon), 1e-8 (nut)
    assert rel(k, d["correct_k"]) < 1e-6
    assert rel(eps, d["correct_epsilon"]) < 1e-6
    assert rel(nut, d["correct_nut"]) < 1e-6
    wall = np.zeros(mesh.nCells, bool)
    wall[mesh.bFaceCell[isWall == 1]] = True
    assert rel(eps[wall], d["correct_epsilon"][wall]) < 1e-9


def test_wall_function_closed_forms():
    kappa, E = 0.41, 9.8
    ypl = O.lib().orc_y_plus_lam(kappa, E)
    assert abs(ypl - np.log(E * ypl) / kappa) < 1e-3 * ypl and 11.0 < ypl < 12.0
    Cmu, nu = 0.09, 1e-5
    # a laminar-sublayer face: yPlus below yPlusLam gives 0
    y, kc = 1e-6, 1e-4
    assert Cmu ** 0.25 * y * np.sqrt(kc) / nu < ypl
    assert O.lib().orc_nutk_wall(Cmu, kappa, E, y, kc, nu) == 0.0
    # a log-layer face: nu*(yPlus*kappa/log(E*yPlus) - 1)
    y, kc = 1e-2, 0.5
    yp = Cmu ** 0.25 * y * np.sqrt(kc) / nu
    ref = nu * (yp * kappa / np.log(E * yp) - 1.0)
    assert abs(O.lib().orc_nutk_wall(Cmu, kappa, E, y, kc, nu) - ref) <= 1e-13 * ref


def test_gbynu_closed_forms():
    n = 3
    g = np.zeros((9, n))
    s, a = 3.0, 2.0
    g[1 * 3 + 0, 0] = s                    # dUx/dy = s: pure shear
    g[0, 1], g[4, 1] = a, -a               # dUx/dx = a, dUy/dy = -a: pure extension
    rng = np.random.default_rng(0)
    g[:, 2] = rng.normal(size=9)
    out = O.gbynu(g)
    assert out[0] == s * s and out[1] == 4 * a * a
    t = g[:, 2].reshape(3, 3)
    ref = (t * (t + t.T)).sum() - (2.0 / 3.0) * np.trace(t) ** 2
    assert abs(out[2] - ref) < 1e-13 * abs(ref)


def test_near_wall_distance_orthogonal_box():
    nx, ny, nz = 5, 6, 4
    m = hexmesh.box_mesh(nx, ny, nz, (2.0, 1.0, 0.5), (1.0, 4.0, 3.0), 0.0, seed=0)
    nI = m["nInternal"]
    mesh = O.Mesh(m["points"], np.arange(0, 4 * m["faces"].shape[0] + 1, 4), m["faces"].ravel(), m["owner"],
                  m["neighbour"], [p[0] for p in m["patches"]], [p[1] for p in m["patches"]],
                  [p[2] for p in m["patches"]])
    nB = len(m["owner"]) - nI
    isWall = np.zeros(nB, np.int32)
    plane = {}
    for name, start, size in m["patches"]:
        if name in ("lowerWall", "upperWall", "back", "front"):
            isWall[start - nI:start - nI + size] = 1
            plane[name] = slice(start - nI, start - nI + size)
    y = O.near_wall_dist(mesh, isWall)
    C = mesh.C
    dist = {"lowerWall": lambda c: C[c, 1], "upperWall": lambda c: 1.0 - C[c, 1], "back": lambda c: C[c, 2],
            "front": lambda c: 0.5 - C[c, 2]}
    bfc = mesh.bFaceCell
    touching = {name: set(bfc[s]) for name, s in plane.items()}
    for name, s in plane.items():
        for i in range(s.start, s.stop):
            c = bfc[i]
            ref = min(dist[w](c) for w in touching if c in touching[w])
            assert abs(y[i] - ref) <= 1e-14, (name, i, y[i], ref)
    assert np.all(y[isWall == 0] == 0.0)
