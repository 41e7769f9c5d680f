"""The oracle's scaled coarse correction (brae's corrScaling, device_amg_vcycle.cu:1113-1126).

The HIP library (and the oracle's mode 1) fuses the scaled correction with the first post-sweep,
x + alpha*e + omega*(r - alpha*Ae)/D, instead of the reference's xg += alpha*c followed by a sweep
that forms A xg afresh (the oracle's mode 2, a line-by-line restatement). The two are the same
arithmetic up to rounding: r - alpha*Ae = b - A(x + alpha*e). Pinned here: one V-cycle agrees to
rounding, the PCG takes the same iteration count and reaches the same solution; and alpha itself is
pinned by the property it is for -- with no smoothing, the corrected residual is orthogonal to the
correction (z.r = z.Az for z = alpha*e), which the unscaled cycle with coarse smoothing does not give.
CPU only."""
import numpy as np

import oracle_lib as O
from brae_amd import hexmesh


def system(seed=4, shape=(14, 9, 6)):
    m = hexmesh.box_mesh(*shape, length=(4.0, 1.0, 0.5), grading=(1.0, 20.0, 20.0), jitter=0.1, seed=seed)
    nI, nC = m["nInternal"], m["nCells"]
    own = np.ascontiguousarray(m["owner"][:nI])
    diag, up = np.zeros(nC), np.zeros(nI)
    O.lib().orc_laplacian(nC, nI, O.p(own), O.p(m["neighbour"]), O.p(m["nonOrthDeltaCoeffs"]),
                          O.p(np.random.default_rng(seed).uniform(0.5, 2.0, nI)), O.p(O.f64(m["magSf"][:nI])),
                          O.p(diag), O.p(up), None)
    diag, up = -diag, -up
    diag[0] *= 2.0   # reference cell: definite
    return m, own, diag, up


def test_fused_scaling_matches_reference_order():
    m, own, diag, up = system()
    nC = m["nCells"]
    ag = O.brae_agglomeration(nC, own, m["neighbour"], m["magSf"][:len(own)], target=8)
    assert ag.nLevels >= 3
    r = np.random.default_rng(0).normal(size=nC)
    for pre, post in ((1, 1), (0, 2), (2, 2)):
        z1 = O.brae_vcycle(ag, diag, up, r, nPre=pre, nPost=post, scale=1)
        z2 = O.brae_vcycle(ag, diag, up, r, nPre=pre, nPost=post, scale=2)
        z0 = O.brae_vcycle(ag, diag, up, r, nPre=pre, nPost=post, scale=0)
        assert np.abs(z1 - z2).max() < 1e-11 * np.abs(z2).max(), (pre, post)
        assert np.abs(z0 - z2).max() > 1e-4 * np.abs(z2).max()   # the scaling does something here
    b = np.random.default_rng(1).normal(size=nC)
    x1, p1 = O.brae_amgpcg(ag, diag, up, b, np.zeros(nC), 1e-10, flexible=True, scale=1)
    x2, p2 = O.brae_amgpcg(ag, diag, up, b, np.zeros(nC), 1e-10, flexible=True, scale=2)
    assert p1.nIterations == p2.nIterations and p1.finalResidual < 1e-10
    assert np.abs(x1 - x2).max() < 1e-8 * np.abs(x2).max()


def test_scaled_correction_is_orthogonal():
    m, own, diag, up = system()
    nC = m["nCells"]
    ag = O.brae_agglomeration(nC, own, m["neighbour"], m["magSf"][:len(own)], target=8)
    r = np.random.default_rng(2).normal(size=nC)
    amul = lambda v: O.amul(nC, own, m["neighbour"], diag, up, None, v)
    # no finest post-sweep but smoothing below: V(1,0) at the finest is x0 + alpha*e with x0 = omega*r/D
    z = O.brae_vcycle(ag, diag, up, r, nPre=1, nPost=0, scale=1)
    x0 = 0.8 * r / diag
    e = z - x0
    rr = r - amul(x0)
    assert abs(e @ rr - e @ amul(e)) < 1e-9 * abs(e @ rr)
    zu = O.brae_vcycle(ag, diag, up, r, nPre=1, nPost=0, scale=0)
    eu = zu - x0
    assert abs(eu @ rr - eu @ amul(eu)) > 1e-3 * abs(eu @ rr)
