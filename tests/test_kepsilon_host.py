"""The library's host near-wall distance (mfv_near_wall_dist) against the oracle's restatement of the
reference's nearWallDist, bit for bit: on pitzDaily (the step's corners, where a neighbouring wall face is
closer than the cell's own) and on a graded, jittered channel whose four walls meet at corners. Host code
only: runs without a GPU."""
import numpy as np

import oracle_lib as O
import test_oracle_kepsilon as T
from brae_amd import hexmesh, mfv


def test_near_wall_distance_pitzdaily_bit_exact():
    d, mesh, U, Ub, uT, kT, kV, eT, eV, k, kb, eps, epsb, nut, nutb, isWall = T.pitz("corr")
    ref = O.near_wall_dist(mesh, isWall)
    got = mfv.near_wall_dist(mesh.nInternal, mesh.points, mesh.faceOffsets, mesh.faceVerts, mesh.bFaceCell, mesh.C,
                             isWall)
    assert np.array_equal(got, ref)
    # the corners matter: some wall faces take a neighbouring face's distance, not 1/deltaCoeffs
    w = isWall == 1
    assert np.any(got[w] < 1.0 / mesh.bDeltaCoeffs[w] * (1 - 1e-9))


def test_near_wall_distance_channel_bit_exact():
    m = hexmesh.box_mesh(12, 9, 7, (4.0, 1.0, 0.5), (1.0, 20.0, 20.0), 0.1, seed=5)
    nI = m["nInternal"]
    mesh = O.Mesh(m["points"], np.arange(0, 4 * m["faces"].shape[0] + 1, 4), m["faces"].ravel(), m["owner"],
                  m["neighbour"])
    nB = len(m["owner"]) - nI
    isWall = np.zeros(nB, np.int32)
    for name, start, size in m["patches"]:
        if name in ("lowerWall", "upperWall", "back", "front"):
            isWall[start - nI:start - nI + size] = 1
    ref = O.near_wall_dist(mesh, isWall)
    got = mfv.near_wall_dist(nI, mesh.points, mesh.faceOffsets, mesh.faceVerts, mesh.bFaceCell, mesh.C, isWall)
    assert np.array_equal(got, ref)
    assert np.all(got[isWall == 1] > 0) and np.all(got[isWall == 0] == 0)
