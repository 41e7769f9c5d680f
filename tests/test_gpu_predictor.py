"""GPU parity of the pressure predictor's momentum glue (rAU, H(), HbyA, constrainHbyA, phiHbyA,
adjustPhi) against the oracle: bit-exact, the adjustPhi sums in the context's sequential mode; with the
default tree sums the mass correction agrees to rounding and continuity closes."""
import ctypes

import numpy as np
import pytest

import gpu_util as G
import oracle_lib as O
from brae_amd import hexmesh, mfv
from brae_amd.mfv import ptr

pytestmark = pytest.mark.gpu


def permuted(m, perm):
    """The box renumbered: cell c becomes perm[c]. Internal faces are re-oriented owner < neighbour (Sf and
    the correction vector change sign, the weight becomes 1 - w) and re-sorted into upper-triangular order;
    boundary faces keep their order. A random permutation leaves a band close to the cell count that does
    not divide it, so the plane-slab gathers meet a partial last plane."""
    nI, nC = m["nInternal"], m["nCells"]
    o, n = perm[m["owner"][:nI]], perm[m["neighbour"]]
    flip = o > n
    lo, hi = np.where(flip, n, o), np.where(flip, o, n)
    order = np.lexsort((hi, lo))
    out = dict(m)
    out["owner"] = np.concatenate([lo[order], perm[m["owner"][nI:]]]).astype(np.int32)
    out["neighbour"] = hi[order].astype(np.int32)
    sgn = np.where(flip, -1.0, 1.0)[order]
    for k in ("Sf", "corrVec"):
        a = np.array(m[k], dtype=float)
        a[:nI] = a[:nI][order] * sgn[:, None]
        out[k] = a
    for k in ("Cf",):
        a = np.array(m[k], dtype=float)
        a[:nI] = a[:nI][order]
        out[k] = a
    w = np.array(m["weights"], dtype=float)
    w[:nI] = np.where(flip, 1.0 - w[:nI], w[:nI])[order]
    out["weights"] = w
    for k in ("magSf", "deltaCoeffs", "nonOrthDeltaCoeffs"):
        a = np.array(m[k], dtype=float)
        a[:nI] = a[:nI][order]
        out[k] = a
    inv = np.empty(nC, np.int64)
    inv[perm] = np.arange(nC)
    out["C"] = np.asarray(m["C"])[inv]
    out["V"] = np.asarray(m["V"])[inv]
    return out


class Case:
    """An asymmetric momentum matrix (upwind convection + laplacian) on a jittered hex box, boundary
    coefficients per face and component, sources, a velocity, host and device copies."""

    def __init__(self, nx, ny, nz, seed, permute=False):
        m = hexmesh.box_mesh(nx, ny, nz, length=(2.0, 1.0, 0.5), grading=(1.0, 4.0, 4.0), jitter=0.1, seed=seed)
        if permute:
            m = permuted(m, np.random.default_rng(seed + 100).permutation(m["nCells"]))
        nI, nC = m["nInternal"], m["nCells"]
        self.m, self.nI, self.nC = m, nI, nC
        self.own, self.nei, self.bfc = O.i32(m["owner"][:nI]), O.i32(m["neighbour"]), O.i32(m["owner"][nI:])
        nB = self.nB = len(self.bfc)
        rng = np.random.default_rng(seed)
        phi = rng.normal(size=nI) * m["magSf"][:nI]
        dU, uU, lU = np.zeros(nC), np.zeros(nI), np.zeros(nI)
        O.lib().orc_div_upwind(nC, nI, O.p(self.own), O.p(self.nei), O.p(O.f64(phi)), O.p(dU), O.p(uU), O.p(lU))
        dL, uL, lL = np.zeros(nC), np.zeros(nI), np.zeros(nI)
        O.lib().orc_laplacian(nC, nI, O.p(self.own), O.p(self.nei), O.p(O.f64(m["deltaCoeffs"])),
                              O.p(np.full(nI, 1e-2)), O.p(O.f64(m["magSf"][:nI])), O.p(dL), O.p(uL), O.p(lL))
        self.diag = dU - dL + 2.0 * (np.abs(uU).max() if nI else 1.0)
        self.upper, self.lower = uU - uL, lU - lL
        self.ic = [rng.uniform(0.1, 1.0, nB) for _ in range(3)]
        self.bc = [rng.normal(size=nB) for _ in range(3)]
        self.src = [rng.normal(size=nC) * m["V"] for _ in range(3)]
        self.U = [rng.normal(size=nC) for _ in range(3)]
        self.Ub = [rng.normal(size=nB) for _ in range(3)]
        self.takeU = (rng.uniform(size=nB) < 0.4).astype(np.int32)
        self.V, self.w, self.Sf = O.f64(m["V"]), O.f64(m["weights"]), O.f64(m["Sf"])
        self.case = G.Case(nC, self.own, self.nei, self.bfc)
        self.d = {k: G.dev(v) for k, v in dict(diag=self.diag, upper=self.upper, lower=self.lower, V=self.V,
                                                 w=self.w, Sf=self.Sf.ravel()).items()}
        self.dic = [G.dev(a) for a in self.ic]

    def call(self, name, *args):
        self.case.call(name, self.case.mesh.h, *args)

    def ldu(self):
        return mfv.ldu(self.d["diag"], self.d["upper"], self.d["lower"])


# one cell (no internal faces), a ragged 3x2x1 box, a small and a 52 k-cell box (plane slabs)
# and the 23 k-cell box renumbered at random (a band of nearly the cell count that does not divide it)
@pytest.fixture(scope="module", params=[(1, 1, 1, 5), (3, 2, 1, 6), (9, 7, 5, 3), (61, 37, 23, 11),
                                        (41, 29, 19, 12, True)],
                ids=["1x1x1", "3x2x1", "9x7x5", "61x37x23", "41x29x19-permuted"])
def case(request):
    return Case(*request.param)


def test_rau_and_hbya_bit_exact(case):
    C = case
    print("case:", C.nC, flush=True)
    D_ref = O.fold_boundary_diag(C.nC, C.bfc, *C.ic, C.diag)
    rAU_ref = O.reciprocal_v(D_ref, C.V)
    dD, dR = G.zeros(C.nC), G.zeros(C.nC)
    C.call("mfv_fold_boundary_diag", ptr(C.d["diag"]), *(ptr(a) for a in C.dic), ptr(dD))
    C.call("mfv_reciprocal_v", ptr(dD), ptr(C.d["V"]), ptr(dR))
    assert np.array_equal(G.host(dD), D_ref)
    assert np.array_equal(G.host(dR), rAU_ref)
    A = C.ldu()
    for k in range(3):
        bd_ref = O.boundary_diag_term(*C.ic, k)
        dbd = G.zeros(C.nB)
        C.call("mfv_boundary_diag_term", *(ptr(a) for a in C.dic), k, ptr(dbd))
        assert np.array_equal(G.host(dbd), bd_ref), k
        H_ref = O.matrix_h(C.nC, C.own, C.nei, C.upper, C.lower, C.U[k], C.src[k], C.bfc, bd_ref, C.bc[k], C.V)
        dU, dS, dB, dH, dHb = G.dev(C.U[k]), G.dev(C.src[k]), G.dev(C.bc[k]), G.zeros(C.nC), G.zeros(C.nC)
        C.call("mfv_matrix_h", ctypes.byref(A), ptr(dU), ptr(dS), ptr(dbd), ptr(dB), ptr(C.d["V"]), None, 1, ptr(dH))
        C.call("mfv_matrix_h", ctypes.byref(A), ptr(dU), ptr(dS), ptr(dbd), ptr(dB), ptr(C.d["V"]), ptr(dR), 1,
               ptr(dHb))
        assert np.array_equal(G.host(dH), H_ref), k
        assert np.array_equal(G.host(dHb), rAU_ref * H_ref), k
    # a knocked-out component comes back as zeros
    dZ = G.dev(np.ones(C.nC))
    C.call("mfv_matrix_h", ctypes.byref(A), ptr(dU), ptr(dS), ptr(dbd), ptr(dB), ptr(C.d["V"]), ptr(dR), 0, ptr(dZ))
    assert not G.host(dZ).any()


def test_three_components_in_one_pass(case):
    """mfv_matrix_h3 gives three mfv_matrix_h calls' values (and the oracle's), with and without rAU, and
    zeros for a knocked-out component"""
    C = case
    print("case:", C.nC, flush=True)
    A = C.ldu()
    rAU = O.reciprocal_v(O.fold_boundary_diag(C.nC, C.bfc, *C.ic, C.diag), C.V)
    dR = G.dev(rAU)
    bd = [O.boundary_diag_term(*C.ic, k) for k in range(3)]
    dPsi, dSrc = [G.dev(u) for u in C.U], [G.dev(s) for s in C.src]
    dBd, dBs = [G.dev(b) for b in bd], [G.dev(b) for b in C.bc]
    for r in (None, dR):
        for valid in (None, [1, 0, 1]):
            outs = [G.dev(np.full(C.nC, 7.0)) for _ in range(3)]
            v = None if valid is None else (ctypes.c_int * 3)(*valid)
            C.call("mfv_matrix_h3", ctypes.byref(A), mfv.ptr3(dPsi), mfv.ptr3(dSrc), mfv.ptr3(dBd), mfv.ptr3(dBs),
                   ptr(C.d["V"]), None if r is None else ptr(r), v, mfv.ptr3(outs))
            for k in range(3):
                got = G.host(outs[k])
                if valid is not None and not valid[k]:
                    assert not got.any(), k
                    continue
                H = O.matrix_h(C.nC, C.own, C.nei, C.upper, C.lower, C.U[k], C.src[k], C.bfc, bd[k], C.bc[k], C.V)
                ref = H if r is None else rAU * H
                single = G.zeros(C.nC)
                C.call("mfv_matrix_h", ctypes.byref(A), ptr(dPsi[k]), ptr(dSrc[k]), ptr(dBd[k]), ptr(dBs[k]),
                       ptr(C.d["V"]), None if r is None else ptr(r), 1, ptr(single))
                assert np.array_equal(got, ref), k
                assert np.array_equal(got, G.host(single)), k


def test_constrain_and_flux_bit_exact(case):
    C = case
    print("case:", C.nC, flush=True)
    HbyA = [np.random.default_rng(20 + k).normal(size=C.nC) for k in range(3)]
    dHbyA = [G.dev(h) for h in HbyA]
    dTake = G.dev(C.takeU)
    Hb = []
    for k in range(3):
        ref = np.where(C.takeU != 0, C.Ub[k], HbyA[k][C.bfc])
        out = G.zeros(C.nB)
        C.call("mfv_constrain_hbya", ptr(dTake), ptr(G.dev(C.Ub[k])), ptr(dHbyA[k]), ptr(out))
        assert np.array_equal(G.host(out), ref), k
        Hb.append(ref)
    phi_ref = O.vector_flux(C.own, C.nei, C.w, C.Sf[:C.nI], *HbyA)
    dphi = G.zeros(C.nI)
    C.call("mfv_vector_flux", ptr(C.d["w"]), ptr(C.d["Sf"]), *(ptr(h) for h in dHbyA), ptr(dphi))
    assert np.array_equal(G.host(dphi), phi_ref)
    empty = np.zeros(C.nB, np.int32)
    empty[::5] = 1
    for e in (None, empty):
        phiB_ref = O.boundary_flux(C.Sf[C.nI:], *Hb, isEmpty=e)
        dpb = G.zeros(C.nB)
        C.call("mfv_boundary_flux", ptr(C.d["Sf"]), *(ptr(G.dev(h)) for h in Hb),
               None if e is None else ptr(G.dev(e)), ptr(dpb))
        assert np.array_equal(G.host(dpb), phiB_ref)


def adjust(C, phiInt, adjustable, phiB):
    dpb = G.dev(phiB)
    mc, closed = ctypes.c_double(), ctypes.c_int()
    rc = mfv.lib().mfv_adjust_phi(C.case.ctx.h, len(phiB), ptr(G.dev(adjustable)), ptr(dpb), len(phiInt),
                                  ptr(G.dev(phiInt)), ctypes.byref(mc), ctypes.byref(closed))
    return rc, mc.value, bool(closed.value), G.host(dpb)


def test_adjust_phi(case):
    C = case
    print("case:", C.nC, flush=True)
    rng = np.random.default_rng(5)
    phiInt = rng.normal(size=C.nI)
    phiB = rng.normal(size=C.nB)
    adj = (rng.uniform(size=C.nB) < 0.3).astype(np.int32)
    if C.nB < 50:
        adj[:] = 1                                       # a few faces: every one adjustable
    rc_ref, mc_ref, closed_ref, out_ref = O.adjust_phi(phiInt, adj, phiB)
    assert rc_ref == 0
    # the host loops' order: bit-exact
    C.case.ctx.sequential_reduce(True)
    try:
        rc, mc, closed, out = adjust(C, phiInt, adj, phiB)
    finally:
        C.case.ctx.sequential_reduce(False)
    assert rc == 0 and mc == mc_ref and closed == closed_ref
    assert np.array_equal(out, out_ref)
    # tree sums: the correction to rounding, continuity closed
    rc, mc, closed, out = adjust(C, phiInt, adj, phiB)
    assert rc == 0 and abs(mc - mc_ref) <= 1e-12 * abs(mc_ref)
    assert abs(out.sum()) <= 1e-11 * np.abs(phiB).sum()
    # no adjustable outflow and an imbalance: refused, the flux untouched
    none = np.zeros(C.nB, np.int32)
    assert O.adjust_phi(phiInt, none, phiB)[0] == 1
    rc, _, _, out = adjust(C, phiInt, none, phiB)
    assert rc == mfv.MFV_ERR_INVALID and np.array_equal(out, phiB)
    assert "continuity" in mfv.lib().mfv_last_error(C.case.ctx.h).decode()


def test_relax_bit_exact(case):
    """fvMatrix::relax (the momentum matrix's diagonal, and the three sources in the same pass) and its
    component reductions: the oracle's values to the bit"""
    C = case
    print("case:", C.nC, flush=True)
    ic = [a.copy() for a in C.ic]
    ic[1][::3] *= -1.5                                   # slip-like faces: the components differ
    dic = [G.dev(a) for a in ic]
    mag_ref, mn_ref = O.cmpt_mag_max_min(*ic)
    dmag, dmn = G.zeros(C.nB), G.zeros(C.nB)
    lib = mfv.lib()
    assert lib.mfv_cmpt_max_mag3(C.case.ctx.h, C.nB, *(ptr(a) for a in dic), ptr(dmag)) == 0
    assert lib.mfv_cmpt_min3(C.case.ctx.h, C.nB, *(ptr(a) for a in dic), ptr(dmn)) == 0
    assert np.array_equal(G.host(dmag), mag_ref) and np.array_equal(G.host(dmn), mn_ref)
    diag = C.diag.copy()
    diag[::4] *= 0.05                                    # rows where the off-diagonal bound wins
    A = mfv.ldu(G.dev(diag), C.d["upper"], C.d["lower"])
    cyc = np.zeros(C.nC)
    cyc[::17] = 3.0
    for alpha, cy in ((0.7, None), (0.3, cyc), (1.0, None)):
        d_ref, delta_ref, src_ref = O.relax_matrix(C.nC, C.own, C.nei, C.upper, C.lower, C.bfc, mag_ref, mn_ref, alpha,
                                                   diag, psi=C.U, source=C.src, cycSumOff=cy)
        dR, dD = G.zeros(C.nC), G.zeros(C.nC)
        dS = [G.dev(s) for s in C.src]
        dP = [G.dev(u) for u in C.U]
        C.call("mfv_relax_diag", ctypes.byref(A), ptr(dic[0]), alpha, None if cy is None else ptr(G.dev(cy)),
               ptr(dmag), ptr(dmn), ptr(dR), ptr(dD), 3, mfv.ptr3(dP), mfv.ptr3(dS))
        assert np.array_equal(G.host(dR), d_ref), alpha
        assert np.array_equal(G.host(dD), delta_ref), alpha
        for k in range(3):
            assert np.array_equal(G.host(dS[k]), src_ref[k]), (alpha, k)
    # the reference's scalar form: one internalCoeffs array, |iC| added and iC removed; no source update
    d_ref, delta_ref, _ = O.relax_matrix(C.nC, C.own, C.nei, C.upper, C.lower, C.bfc, np.abs(C.ic[0]), C.ic[0], 0.5,
                                         diag)
    dR = G.zeros(C.nC)
    C.call("mfv_relax_diag", ctypes.byref(A), ptr(C.dic[0]), 0.5, None, None, None, ptr(dR), None, 0, None, None)
    assert np.array_equal(G.host(dR), d_ref)
    # alpha <= 0 is refused by name (the host skips the relaxation)
    rc = lib.mfv_relax_diag(C.case.ctx.h, C.case.mesh.h, ctypes.byref(A), ptr(C.dic[0]), 0.0, None, None, None,
                            ptr(dR), None, 0, None, None)
    assert rc == mfv.MFV_ERR_INVALID and "alpha" in lib.mfv_last_error(C.case.ctx.h).decode()


def test_relax_field(case):
    C = case
    rng = np.random.default_rng(12)
    p, prev = rng.normal(size=C.nC), rng.normal(size=C.nC)
    for alpha in (0.3, 0.7, 1.0, 0.0):
        dp = G.dev(p)
        assert mfv.lib().mfv_relax_field(C.case.ctx.h, C.nC, ptr(dp), ptr(G.dev(prev)), alpha) == 0
        assert np.array_equal(G.host(dp), O.relax_field(p, prev, alpha)), alpha


def test_linear_upwind_corr_bit_exact(case):
    """linearUpwind's deferred correction: three components in one pass and one at a time, the oracle's
    face-order sums to the bit (zero fluxes included: the neighbour is the upwind cell there)"""
    C = case
    print("case:", C.nC, flush=True)
    rng = np.random.default_rng(31)
    phi = rng.normal(size=C.nI) * C.m["magSf"][:C.nI]
    phi[::7] = 0.0
    grads = [rng.normal(size=(C.nC, 3)) for _ in range(3)]
    Cf, Cc = O.f64(C.m["Cf"]), O.f64(C.m["C"])
    dphi, dCf, dC = G.dev(phi), G.dev(Cf.ravel()), G.dev(Cc.ravel())
    dg = [G.dev(g.ravel()) for g in grads]
    refs = [O.linear_upwind_corr(C.nC, C.own, C.nei, phi, Cf[:C.nI], Cc, g) for g in grads]
    outs = [G.dev(np.full(C.nC, 9.0)) for _ in range(3)]
    C.call("mfv_linear_upwind_corr", ptr(dphi), ptr(dCf), ptr(dC), 3, mfv.ptr3(dg), mfv.ptr3(outs))
    for k in range(3):
        assert np.array_equal(G.host(outs[k]), refs[k]), k
        one = G.zeros(C.nC)
        C.call("mfv_linear_upwind_corr", ptr(dphi), ptr(dCf), ptr(dC), 1, (ctypes.c_void_p * 1)(dg[k].data_ptr()),
               (ctypes.c_void_p * 1)(one.data_ptr()))
        assert np.array_equal(G.host(one), refs[k]), k
    if C.nI == 0:
        # a mesh without internal faces: phi, Cf and C may be null and the correction is zero
        zs = [G.dev(np.full(C.nC, 9.0)) for _ in range(3)]
        C.call("mfv_linear_upwind_corr", None, None, None, 3, mfv.ptr3(dg), mfv.ptr3(zs))
        for k in range(3):
            assert np.array_equal(G.host(zs[k]), np.zeros(C.nC)), k


def test_viscous_stress_bit_exact(case):
    """nu*dev2(T(grad U)) and V*div of the stress tensor (the explicit part of divDevReff), the oracle's
    values to the bit; empty boundary faces dropped"""
    C = case
    print("case:", C.nC, flush=True)
    rng = np.random.default_rng(41)
    gradU = rng.normal(size=(9, C.nC))
    nu = rng.uniform(1e-3, 1e-1, C.nC)
    sig_ref = O.scaled_dev2t(gradU, nu)
    dsig = G.zeros(9 * C.nC)
    assert mfv.lib().mfv_scaled_dev2t(C.case.ctx.h, C.nC, ptr(G.dev(gradU.ravel())), ptr(G.dev(nu)), ptr(dsig)) == 0
    assert np.array_equal(G.host(dsig).reshape(9, C.nC), sig_ref)
    Tb = rng.normal(size=(9, C.nB))
    empty = np.zeros(C.nB, np.int32)
    empty[::4] = 1
    for e in (None, empty):
        ref = O.tensor_div(C.nC, C.own, C.nei, C.w, C.Sf, C.bfc, sig_ref, Tb, isEmpty=e)
        out = [G.dev(np.full(C.nC, 5.0)) for _ in range(3)]
        C.call("mfv_tensor_div_source", ptr(C.d["w"]), ptr(C.d["Sf"]), None if e is None else ptr(G.dev(e)), ptr(dsig),
               ptr(G.dev(Tb.ravel())), *(ptr(o) for o in out))
        for k in range(3):
            assert np.array_equal(G.host(out[k]), ref[k]), k
        # the fused pass forms each cell's stress from grad(U) where it is read: the same bits
        fused = [G.dev(np.full(C.nC, 5.0)) for _ in range(3)]
        C.call("mfv_div_dev_reff_explicit", ptr(C.d["w"]), ptr(C.d["Sf"]), None if e is None else ptr(G.dev(e)),
               ptr(G.dev(gradU.ravel())), ptr(G.dev(nu)), ptr(G.dev(Tb.ravel())), *(ptr(o) for o in fused))
        for k in range(3):
            assert np.array_equal(G.host(fused[k]), ref[k]), k
