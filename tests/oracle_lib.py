"""ctypes binding of the oracle (oracle/liboracle.so) for the tests, smoke() and bench.py's CPU
baseline. Test infrastructure: the product never imports this module."""
import collections
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
_LIB = None


class Perf(ctypes.Structure):
    _fields_ = [("initialResidual", ctypes.c_double), ("finalResidual", ctypes.c_double),
                ("nIterations", ctypes.c_int)]


class GamgControls(ctypes.Structure):
    _fields_ = [("smoother", ctypes.c_int), ("tolerance", ctypes.c_double), ("relTol", ctypes.c_double),
                ("maxIter", ctypes.c_int), ("minIter", ctypes.c_int),
                ("nPreSweeps", ctypes.c_int), ("preSweepsLevelMultiplier", ctypes.c_int),
                ("maxPreSweeps", ctypes.c_int), ("nPostSweeps", ctypes.c_int),
                ("postSweepsLevelMultiplier", ctypes.c_int), ("maxPostSweeps", ctypes.c_int),
                ("nFinestSweeps", ctypes.c_int), ("scaleCorrection", ctypes.c_int), ("nVcycles", ctypes.c_int)]


SMOOTHERS = {"GaussSeidel": 0, "symGaussSeidel": 1, "DIC": 2, "DICGaussSeidel": 3}
BC_ZERO_GRADIENT, BC_FIXED_VALUE, BC_EMPTY = 0, 1, 2


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(ORACLE_DIR, "liboracle.so")
        if not os.path.exists(path):
            subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        I, D = ctypes.c_int, ctypes.c_double
        sig = {
            "orc_geometry": (I, [I, P, I, P, P, I, I, P, P] + [P] * 9),
            "orc_patch_delta": (None, [I, P, P, P, P, P, P]),
            "orc_ldu_addressing": (I, [I, I, P, P, P, P, P]),
            "orc_laplacian": (None, [I, I, P, P, P, P, P, P, P, P]),
            "orc_laplacian_bcoeffs": (None, [I, P, P, P, P, P, P, P]),
            "orc_div_upwind": (None, [I, I, P, P, P, P, P, P]),
            "orc_div_bcoeffs": (None, [I, P, P, P, P, P]),
            "orc_div_linear": (None, [I, I, P, P, P, P, P, P, P]),
            "orc_laplacian_corr_limited": (None, [I, I, P, P, P, P, P, P, P, P, P, D, P, P]),
            "orc_fold": (None, [I, P, P, I, P, P, P, P, P, P]),
            "orc_set_reference": (None, [P, P, I, D]),
            "orc_amul": (None, [I, I, P, P, P, P, P, P, P]),
            "orc_residual": (None, [I, I, P, P, P, P, P, P, P, P]),
            "orc_norm_factor": (D, [I, I, P, P, P, P, P, P, P]),
            "orc_sum_mag": (D, [I, P]),
            "orc_sum_prod": (D, [I, P, P]),
            "orc_face_flux": (None, [I, P, P, P, P, P, P]),
            "orc_fold_boundary_diag": (None, [I, I, P, P, P, P, P, P]),
            "orc_reciprocal_v": (None, [I, P, P, P]),
            "orc_boundary_diag_term": (None, [I, P, P, P, I, P]),
            "orc_matrix_h": (None, [I, I, P, P, P, P, P, P, I, P, P, P, P, P]),
            "orc_vector_flux": (None, [I, P, P, P, P, P, P, P, P]),
            "orc_boundary_flux": (None, [I, P, P, P, P, P, P]),
            "orc_adjust_phi": (I, [I, P, I, P, P, P, P]),
            "orc_relax_matrix": (None, [I, I, P, P, P, P, I, P, P, P, P, D, P, I, P, P, P]),
            "orc_cmpt_mag_max_min": (None, [I, P, P, P, P, P]),
            "orc_relax_field": (None, [I, P, P, D]),
            "orc_linear_upwind_corr": (None, [I, I, P, P, P, P, P, P, P]),
            "orc_scaled_dev2t": (None, [I, P, P, P]),
            "orc_tensor_div": (None, [I, I, P, P, P, P, I, P, P, P, P, P, P, P]),
            "orc_grad_gauss": (None, [I, I, P, P, I, P, P, P, P, P, P, P, P]),
            "orc_gauss_seidel": (None, [I, I, P, P, P, P, P, P, I]),
            "orc_sym_gauss_seidel": (None, [I, I, P, P, P, P, P, P, I]),
            "orc_dic_precondition": (None, [I, I, P, P, P, P, P, P]),
            "orc_laplacian_corr": (None, [I, I, P, P, P, P, P, P, P, P, P]),
            "orc_flux_correct": (None, [I, P, P, P, P, P, P, P, P]),
            "orc_pcg_dic": (Perf, [I, I, P, P, P, P, P, P, D, D, I, I]),
            "orc_faceareapair_weights": (None, [I, P, P, P]),
            "orc_of_agglomeration": (P, [I, I, P, P, P, I, P]),
            "orc_brae_agglomeration": (P, [I, I, P, P, P, I]),
            "orc_aggl_nlevels": (I, [P]),
            "orc_aggl_sizes": (None, [P, I, P, P, P, P]),
            "orc_aggl_level": (None, [P, I, P, P, P, P, P]),
            "orc_aggl_free": (None, [P]),
            "orc_galerkin_of": (None, [P, P, P, P, P]),
            "orc_galerkin_brae": (None, [P, P, P, P, P]),
            "orc_gamg_default_controls": (None, [P]),
            "orc_gamg_solve": (Perf, [P, P, P, P, P, P]),
            "orc_pcg_gamg_solve": (Perf, [P, P, P, P, P, D, D, I, I, P]),
            "orc_brae_amgpcg": (Perf, [P, P, P, P, P, D, D, I, I, I, I, I, I, I, D, P]),
            "orc_brae_vcycle": (None, [P, P, P, P, P, I, I, I, I]),
            "orc_pbicgstab": (Perf, [I, I, P, P, P, P, P, P, I, P, P, P, P, D, D, I, I]),
            "orc_gamg_ref": (Perf, [I, I, P, P, P, P, P, P, I, P, P, P, P, P, D, D, I, I, I, I, I, P]),
            "orc_gamg_ref_pair": (I, [I, I, P, P, P, I, P]),
            "orc_gamg_ref_coarsen": (I, [I, I, P, P, P, P, P, P, P, I, P, P, P, P, P, P]),
            "orc_gamg_ref_gauss_seidel": (None, [I, I, P, P, P, P, P, P, P, I]),
            "orc_gamg_ref_dense": (None, [I, I, P, P, P, P, P, P, P]),
            "orc_gamg_ref_scale": (None, [I, I, P, P, P, P, P, P, P]),
            "orc_grad_u_boundary": (None, [I, I, I, P, P, P, P, P, P, P, P, P, P, P, P, P]),
            "orc_simple_step": (I, [P, P, P]),
            "orc_kepsilon_default_coeffs": (None, [P]),
            "orc_near_wall_dist": (None, [I, I, P, P, P, P, P, P, P]),
            "orc_y_plus_lam": (D, [D, D]),
            "orc_nutk_wall": (D, [D, D, D, D, D, D]),
            "orc_gbynu": (None, [I, P, P]),
            "orc_kepsilon_eqns": (None, [P] + [P] * 10 + [P] * 12),
            "orc_kepsilon_correct": (I, [P, P, P, P, P, P, P, P, P, P, P, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


_KEEP = collections.deque(maxlen=256)


def p(a):
    """data pointer of a C-contiguous numpy array (None passes NULL). The array is kept alive for
    the next calls, so a temporary built in an argument list outlives the call that uses it."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    _KEEP.append(a)
    return a.ctypes.data


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Mesh:
    """A polyMesh (points, faces, owner, neighbour, patches) with the oracle's geometry."""

    def __init__(self, points, faceOffsets, faceVerts, owner, neighbour, patchNames=(), patchStart=(),
                 patchSize=(), patchTypes=()):
        self.points = f64(points)
        self.faceOffsets = i32(faceOffsets)
        self.faceVerts = i32(faceVerts)
        self.owner = i32(owner)
        self.neighbour = i32(neighbour)
        self.nFaces = self.owner.size
        self.nInternal = self.neighbour.size
        self.nCells = int(max(self.owner.max(), self.neighbour.max() if self.nInternal else 0)) + 1
        self.patchNames = [str(x) for x in patchNames]
        self.patchStart = [int(x) for x in patchStart]
        self.patchSize = [int(x) for x in patchSize]
        self.patchTypes = [str(x) for x in patchTypes]
        self.geometry()

    @classmethod
    def from_npz(cls, d):
        return cls(d["points"], d["faceOffsets"], d["faceVerts"], d["owner"], d["neighbour"],
                   d["patchNames"], d["patchStart"], d["patchSize"], d["patchTypes"])

    def geometry(self):
        nF, nI, nC = self.nFaces, self.nInternal, self.nCells
        self.Cf = np.zeros((nF, 3)); self.Sf = np.zeros((nF, 3)); self.magSf = np.zeros(nF)
        self.C = np.zeros((nC, 3)); self.V = np.zeros(nC)
        self.weights = np.zeros(nI); self.deltaCoeffs = np.zeros(nI)
        self.nonOrthDeltaCoeffs = np.zeros(nI); self.corrVec = np.zeros((nI, 3))
        lib().orc_geometry(self.points.shape[0], p(self.points), nF, p(self.faceOffsets), p(self.faceVerts),
                           nC, nI, p(self.owner), p(self.neighbour), p(self.Cf), p(self.Sf), p(self.magSf),
                           p(self.C), p(self.V), p(self.weights), p(self.deltaCoeffs),
                           p(self.nonOrthDeltaCoeffs), p(self.corrVec))
        nB = nF - nI
        self.bFaceCell = i32(self.owner[nI:])
        self.bDeltaCoeffs = np.zeros(nB)
        lib().orc_patch_delta(nB, p(self.bFaceCell), p(f64(self.Cf[nI:])), p(f64(self.Sf[nI:])),
                              p(f64(self.magSf[nI:])), p(self.C), p(self.bDeltaCoeffs))

    @property
    def nBFaces(self):
        return self.nFaces - self.nInternal

    def patch_slice(self, name):
        k = self.patchNames.index(name)
        s = self.patchStart[k] - self.nInternal
        return slice(s, s + self.patchSize[k])

    def own(self):
        return i32(self.owner[:self.nInternal])

    def addressing(self):
        os_ = np.zeros(self.nCells + 1, np.int32)
        ls = np.zeros(self.nCells + 1, np.int32)
        lo = np.zeros(self.nInternal, np.int32)
        rc = lib().orc_ldu_addressing(self.nCells, self.nInternal, p(self.own()), p(self.neighbour),
                                      p(os_), p(ls), p(lo))
        assert rc == 0
        return os_, ls, lo


def laplacian(mesh, dc, gammaf):
    nC, nI = mesh.nCells, mesh.nInternal
    diag, upper, lower = np.zeros(nC), np.zeros(nI), np.zeros(nI)
    lib().orc_laplacian(nC, nI, p(mesh.own()), p(mesh.neighbour), p(f64(dc)), p(f64(gammaf)),
                        p(f64(mesh.magSf[:nI])), p(diag), p(upper), p(lower))
    return diag, upper, lower


def laplacian_bcoeffs(mesh, bcType, bcValue, gammab):
    nB, nI = mesh.nBFaces, mesh.nInternal
    iC, bC = np.zeros(nB), np.zeros(nB)
    lib().orc_laplacian_bcoeffs(nB, p(i32(bcType)), p(f64(bcValue)), p(f64(gammab)), p(f64(mesh.magSf[nI:])),
                                p(mesh.bDeltaCoeffs), p(iC), p(bC))
    return iC, bC


def fold(mesh, diag, source, bcType, iC, bC):
    diagC, b = np.zeros(mesh.nCells), np.zeros(mesh.nCells)
    lib().orc_fold(mesh.nCells, p(f64(diag)), p(f64(source)), mesh.nBFaces, p(mesh.bFaceCell), p(i32(bcType)),
                   p(f64(iC)), p(f64(bC)), p(diagC), p(b))
    return diagC, b


def grad_gauss(nC, own, nei, bFaceCell, bcType, bcValue, Sf, weights, V, psi):
    """Gauss linear gradient (nC x 3); Sf per face, internal then boundary"""
    g = np.zeros((nC, 3))
    lib().orc_grad_gauss(nC, len(own), p(i32(own)), p(i32(nei)), len(bFaceCell), p(i32(bFaceCell)), p(i32(bcType)),
                         p(f64(bcValue)), p(f64(Sf)), p(f64(weights)), p(f64(V)), p(f64(psi)), p(g))
    return g


def laplacian_corr(nC, own, nei, grad, weights, corrVec, gammaf, magSf, source=None):
    """non-orthogonal correction flux; with a source, also source -= div(ffc) (in place, returned)"""
    ffc = np.zeros(len(own))
    src = None if source is None else f64(source).copy()
    lib().orc_laplacian_corr(nC, len(own), p(i32(own)), p(i32(nei)), p(f64(grad)), p(f64(weights)), p(f64(corrVec)),
                             p(f64(gammaf)), p(f64(magSf)), p(ffc), p(src))
    return ffc, src


def laplacian_corr_limited(nC, own, nei, grad, weights, corrVec, gammaf, magSf, nonOrthDc, vf, k, source=None):
    """`limited k corrected` correction flux (and source -= div(ffc) when a source is given)"""
    ffc = np.zeros(len(own))
    src = None if source is None else f64(source).copy()
    lib().orc_laplacian_corr_limited(nC, len(own), p(i32(own)), p(i32(nei)), p(f64(grad)), p(f64(weights)),
                                     p(f64(corrVec)), p(f64(gammaf)), p(f64(magSf)), p(f64(nonOrthDc)), p(f64(vf)),
                                     float(k), p(ffc), p(src))
    return ffc, src


def flux_correct(own, nei, upper, lower, psi, ffc, phiHbyA):
    phi = np.zeros(len(own))
    lib().orc_flux_correct(len(own), p(i32(own)), p(i32(nei)), p(f64(upper)), p(None if lower is None else f64(lower)),
                           p(f64(psi)), p(None if ffc is None else f64(ffc)), p(f64(phiHbyA)), p(phi))
    return phi


def fold_boundary_diag(nC, bfc, icX, icY, icZ, diag):
    D = np.zeros(nC)
    lib().orc_fold_boundary_diag(nC, len(bfc), p(i32(bfc)), p(f64(icX)), p(f64(icY)), p(f64(icZ)), p(f64(diag)), p(D))
    return D


def reciprocal_v(D, V):
    r = np.zeros(len(D))
    lib().orc_reciprocal_v(len(D), p(f64(D)), p(f64(V)), p(r))
    return r


def boundary_diag_term(icX, icY, icZ, k):
    out = np.zeros(len(icX))
    lib().orc_boundary_diag_term(len(icX), p(f64(icX)), p(f64(icY)), p(f64(icZ)), int(k), p(out))
    return out


def matrix_h(nC, own, nei, upper, lower, psi, source, bfc, bdDiag, bdSrc, V):
    H = np.zeros(nC)
    lib().orc_matrix_h(nC, len(own), p(i32(own)), p(i32(nei)), p(f64(upper)), p(f64(lower)), p(f64(psi)),
                       p(f64(source)), len(bfc), p(i32(bfc)), p(f64(bdDiag)), p(f64(bdSrc)), p(f64(V)), p(H))
    return H


def vector_flux(own, nei, weights, Sf, Ux, Uy, Uz):
    phi = np.zeros(len(own))
    lib().orc_vector_flux(len(own), p(i32(own)), p(i32(nei)), p(f64(weights)), p(f64(Sf)), p(f64(Ux)), p(f64(Uy)),
                          p(f64(Uz)), p(phi))
    return phi


def boundary_flux(SfB, uxb, uyb, uzb, isEmpty=None):
    phiB = np.zeros(len(uxb))
    lib().orc_boundary_flux(len(uxb), p(f64(SfB)), p(f64(uxb)), p(f64(uyb)), p(f64(uzb)),
                            p(None if isEmpty is None else i32(isEmpty)), p(phiB))
    return phiB


def adjust_phi(phiInt, adjustable, phiB):
    """(status, massCorr, closedVolume, adjusted phiB): status 1 = continuity error not removable"""
    out = f64(phiB).copy()
    mc, closed = ctypes.c_double(), ctypes.c_int()
    rc = lib().orc_adjust_phi(len(phiInt), p(f64(phiInt)), len(out), p(i32(adjustable)), p(out),
                              ctypes.byref(mc), ctypes.byref(closed))
    return rc, mc.value, bool(closed.value), out


def gauss_seidel(nC, own, nei, diag, upper, b, psi, nSweeps, symmetric=False):
    psi = f64(psi).copy()
    fn = lib().orc_sym_gauss_seidel if symmetric else lib().orc_gauss_seidel
    fn(nC, len(own), p(i32(own)), p(i32(nei)), p(f64(diag)), p(f64(upper)), p(f64(b)), p(psi), nSweeps)
    return psi


def dic_precondition(nC, own, nei, diag, upper, r):
    w = np.zeros(nC)
    lib().orc_dic_precondition(nC, len(own), p(i32(own)), p(i32(nei)), p(f64(diag)), p(f64(upper)), p(f64(r)), p(w))
    return w


def amul(nC, own, nei, diag, upper, lower, psi):
    out = np.zeros(nC)
    lib().orc_amul(nC, len(own), p(i32(own)), p(i32(nei)), p(f64(diag)), p(f64(upper)),
                   p(None if lower is None else f64(lower)), p(f64(psi)), p(out))
    return out


def norm_factor(nC, own, nei, diag, upper, psi, b):
    return lib().orc_norm_factor(nC, len(own), p(i32(own)), p(i32(nei)), p(f64(diag)), p(f64(upper)), None,
                                 p(f64(psi)), p(f64(b)))


def pcg_dic(nC, own, nei, diag, upper, b, psi, tol, relTol=0.0, maxIter=1000, minIter=0):
    psi = f64(psi).copy()
    perf = lib().orc_pcg_dic(nC, len(own), p(i32(own)), p(i32(nei)), p(f64(diag)), p(f64(upper)), p(f64(b)),
                             p(psi), tol, relTol, maxIter, minIter)
    return psi, perf


class Agglomeration:
    def __init__(self, handle):
        self.h = handle
        L = lib()
        self.nLevels = L.orc_aggl_nlevels(self.h)
        self.levels = []
        for i in range(self.nLevels):
            s = [ctypes.c_int() for _ in range(4)]
            L.orc_aggl_sizes(self.h, i, *[ctypes.addressof(x) for x in s])
            nF, nFF, nCo, nCF = (x.value for x in s)
            m, co, cn, fr = (np.zeros(nF, np.int32), np.zeros(nCF, np.int32), np.zeros(nCF, np.int32),
                             np.zeros(nFF, np.int32))
            ff = np.zeros(nFF, np.int8)
            L.orc_aggl_level(self.h, i, p(m), p(co), p(cn), p(fr), p(ff))
            self.levels.append(dict(nFine=nF, nFineFaces=nFF, nCoarse=nCo, nCoarseFaces=nCF, map=m,
                                    cOwn=co, cNei=cn, faceRestrict=fr, faceFlip=ff))

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.orc_aggl_free(self.h)
            self.h = None

    def total_coarse(self):
        return (sum(l["nCoarse"] for l in self.levels), sum(l["nCoarseFaces"] for l in self.levels))

    def galerkin(self, diag, upper, order="brae"):
        tc, tf = self.total_coarse()
        cd, cu = np.zeros(tc), np.zeros(tf)
        fn = lib().orc_galerkin_brae if order == "brae" else lib().orc_galerkin_of
        fn(self.h, p(f64(diag)), p(f64(upper)), p(cd), p(cu))
        out, oc, of = [], 0, 0
        for l in self.levels:
            out.append((cd[oc:oc + l["nCoarse"]], cu[of:of + l["nCoarseFaces"]]))
            oc += l["nCoarse"]; of += l["nCoarseFaces"]
        return out


def of_agglomeration(nC, own, nei, weights, nCellsInCoarsestLevel=10, forward=True):
    fwd = ctypes.c_int(1 if forward else 0)
    h = lib().orc_of_agglomeration(nC, len(own), p(i32(own)), p(i32(nei)), p(f64(weights)),
                                   nCellsInCoarsestLevel, ctypes.addressof(fwd))
    return Agglomeration(h)


def brae_agglomeration(nC, own, nei, weights, target=64):
    h = lib().orc_brae_agglomeration(nC, len(own), p(i32(own)), p(i32(nei)), p(f64(weights)), target)
    return Agglomeration(h)


def faceareapair_weights(Sf, magSf):
    w = np.zeros(len(magSf))
    lib().orc_faceareapair_weights(len(magSf), p(f64(Sf)), p(f64(magSf)), p(w))
    return w


def gamg_controls(**kw):
    c = GamgControls()
    lib().orc_gamg_default_controls(ctypes.addressof(c))
    for k, v in kw.items():
        if k == "smoother" and isinstance(v, str):
            v = SMOOTHERS[v]
        setattr(c, k, v)
    return c


def gamg_solve(aggl, diag, upper, b, psi, **kw):
    c = gamg_controls(**kw)
    psi = f64(psi).copy()
    perf = lib().orc_gamg_solve(aggl.h, p(f64(diag)), p(f64(upper)), p(f64(b)), p(psi), ctypes.addressof(c))
    return psi, perf


def pcg_gamg_solve(aggl, diag, upper, b, psi, tol, relTol=0.0, maxIter=1000, minIter=0, **kw):
    c = gamg_controls(**kw)
    psi = f64(psi).copy()
    perf = lib().orc_pcg_gamg_solve(aggl.h, p(f64(diag)), p(f64(upper)), p(f64(b)), p(psi), tol, relTol,
                                    maxIter, minIter, ctypes.addressof(c))
    return psi, perf


def brae_amgpcg(aggl, diag, upper, b, psi, tol, relTol=0.0, maxIter=1000, minIter=0, nPre=1, nPost=1,
                flexible=False, single=False, normFactor=0.0, trace=False, scale=False):
    psi = f64(psi).copy()
    tr = np.zeros((maxIter + 1, psi.size)) if trace else None
    perf = lib().orc_brae_amgpcg(aggl.h, p(f64(diag)), p(f64(upper)), p(f64(b)), p(psi), tol, relTol, maxIter,
                                 minIter, nPre, nPost, int(flexible), int(single), int(scale), normFactor, p(tr))
    return (psi, perf, tr) if trace else (psi, perf)


def brae_vcycle(aggl, diag, upper, r, nPre=1, nPost=1, single=False, scale=False):
    z = np.zeros(len(r))
    lib().orc_brae_vcycle(aggl.h, p(f64(diag)), p(f64(upper)), p(f64(r)), p(z), nPre, nPost, int(single), int(scale))
    return z


def cmpt_mag_max_min(x, y, z):
    """per boundary face cmptMax(cmptMag(ic)) and cmptMin(ic)"""
    mag, mn = np.zeros(len(x)), np.zeros(len(x))
    lib().orc_cmpt_mag_max_min(len(x), p(f64(x)), p(f64(y)), p(f64(z)), p(mag), p(mn))
    return mag, mn


def relax_matrix(nC, own, nei, upper, lower, bfc, icMag, icMin, alpha, diag, psi=(), source=(), cycSumOff=None):
    """fvMatrix::relax: (relaxed diag, delta, relaxed sources) for the components psi/source"""
    d = f64(diag).copy()
    src = [f64(s).copy() for s in source]
    ps = [f64(x) for x in psi]
    k = len(src)
    PA = ctypes.c_void_p * max(k, 1)
    pp = PA(*[x.ctypes.data for x in ps]) if k else None
    sp = PA(*[x.ctypes.data for x in src]) if k else None
    delta = np.zeros(nC)
    lib().orc_relax_matrix(nC, len(own), p(i32(own)), p(i32(nei)), p(f64(upper)), p(f64(lower)), len(bfc),
                           p(i32(bfc)), p(f64(icMag)), p(f64(icMin)), p(None if cycSumOff is None else f64(cycSumOff)),
                           alpha, p(d), k, ctypes.cast(pp, ctypes.c_void_p) if k else None,
                           ctypes.cast(sp, ctypes.c_void_p) if k else None, p(delta))
    return d, delta, src


def relax_field(pNew, pPrev, alpha):
    out = f64(pNew).copy()
    lib().orc_relax_field(len(out), p(out), p(f64(pPrev)), alpha)
    return out


def linear_upwind_corr(nC, own, nei, phi, Cf, C, grad):
    """linearUpwind's correction per cell for one component; grad [nC, 3]"""
    corr = np.zeros(nC)
    lib().orc_linear_upwind_corr(nC, len(own), p(i32(own)), p(i32(nei)), p(f64(phi)), p(f64(Cf).ravel()),
                                 p(f64(C).ravel()), p(f64(grad).ravel()), p(corr))
    return corr



def scaled_dev2t(gradU, nu):
    """nu*dev2(T(gradU)); gradU [9, n] component-major (row q = i*3 + j is d(U_j)/d(x_i))"""
    g = f64(gradU)
    sig = np.zeros_like(g)
    lib().orc_scaled_dev2t(g.shape[1], p(g), p(f64(nu)), p(sig))
    return sig


def tensor_div(nC, own, nei, w, Sf, bfc, Tcell, Tbnd, isEmpty=None):
    """V*div(T): (dX, dY, dZ); Tcell [9, nC], Tbnd [9, nB]; Sf [nFaces + nB, 3]"""
    d = [np.zeros(nC) for _ in range(3)]
    lib().orc_tensor_div(nC, len(own), p(i32(own)), p(i32(nei)), p(f64(w)), p(f64(Sf).ravel()), len(bfc), p(i32(bfc)),
                         p(None if isEmpty is None else i32(isEmpty)), p(f64(Tcell)), p(f64(Tbnd)), *(p(x) for x in d))
    return d


# ---- one SIMPLE iteration (oracle/simple.c) ----
P_GAMG_REF, P_BRAE_AMGPCG, P_PCG_GAMG = 0, 1, 2
_VP, _I, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double


class SimpleCase(ctypes.Structure):
    _fields_ = [("nCells", _I), ("nFaces", _I), ("nBFaces", _I),
                ("owner", _VP), ("neighbour", _VP), ("bFaceCell", _VP),
                ("Sf", _VP), ("Cf", _VP), ("magSf", _VP), ("C", _VP), ("V", _VP), ("weights", _VP),
                ("deltaCoeffs", _VP), ("nonOrthDeltaCoeffs", _VP), ("corrVec", _VP), ("bDeltaCoeffs", _VP),
                ("uType", _VP), ("uValue", _VP), ("pType", _VP), ("pValue", _VP),
                ("nu", _D), ("nuEff", _VP), ("nuEffB", _VP), ("relaxU", _D), ("relaxP", _D),
                ("momentumPredictor", _I), ("linearUpwind", _I), ("correctedLaplacian", _I), ("nNonOrthCorr", _I),
                ("tolU", _D), ("relTolU", _D), ("maxIterU", _I), ("minIterU", _I),
                ("tolP", _D), ("relTolP", _D), ("maxIterP", _I), ("minIterP", _I),
                ("pSolver", _I), ("pAggl", _VP), ("amgPre", _I), ("amgPost", _I), ("amgFlexible", _I),
                ("amgSingle", _I), ("amgScale", _I), ("pRefCell", _I), ("pRefValue", _D)]


class SimpleFields(ctypes.Structure):
    _fields_ = [(n, _VP) for n in ("Ux", "Uy", "Uz", "Ubx", "Uby", "Ubz", "p", "pb", "phi", "phiB")]


class SimpleResult(ctypes.Structure):
    _fields_ = [("UInitial", _D * 3), ("UFinal", _D * 3), ("UIters", _I * 3), ("pInitial", _D), ("pFinal", _D),
                ("pIters", _I * 8), ("massCorr", _D), ("adjustStatus", _I)]


def bc_codes(types):
    """patch type names to the oracle's codes: fixedValue and noSlip fix a value, empty drops out"""
    m = {"fixedValue": BC_FIXED_VALUE, "noSlip": BC_FIXED_VALUE, "zeroGradient": BC_ZERO_GRADIENT,
         "empty": BC_EMPTY}
    return i32([m[str(t)] for t in types])


def evaluate_boundary(bfc, types, values, cell):
    """patch values after evaluate(): fixedValue keeps its value, zeroGradient the owner cell's, empty 0"""
    out = np.where(types == BC_FIXED_VALUE, values, 0.0)
    zg = types == BC_ZERO_GRADIENT
    out[zg] = cell[bfc[zg]]
    return out


class SimpleState:
    """Host fields of one SIMPLE run: U (3 cell arrays + 3 boundary arrays), p, p's patch values, phi."""

    def __init__(self, U, Ub, p, pb, phi, phiB):
        self.U = [f64(u).copy() for u in U]
        self.Ub = [f64(u).copy() for u in Ub]
        self.p, self.pb, self.phi, self.phiB = f64(p).copy(), f64(pb).copy(), f64(phi).copy(), f64(phiB).copy()

    def copy(self):
        return SimpleState(self.U, self.Ub, self.p, self.pb, self.phi, self.phiB)


def simple_case(mesh, uType, uValue, pType, pValue, nu=1e-5, relaxU=0.7, relaxP=0.3, momentumPredictor=1,
                linearUpwind=0, correctedLaplacian=0, nNonOrthCorr=0, tolU=1e-10, relTolU=0.0, maxIterU=1000,
                minIterU=0, tolP=1e-10, relTolP=0.0, maxIterP=1000, minIterP=0, pSolver=P_GAMG_REF, aggl=None,
                amgPre=1, amgPost=1, amgFlexible=0, amgSingle=0, amgScale=0, pRefCell=-1, pRefValue=0.0,
                nuEff=None, nuEffB=None):
    """The oracle's case struct for an O.Mesh (its arrays are kept alive on the returned object)."""
    keep = dict(own=mesh.own(), nei=i32(mesh.neighbour), bfc=i32(mesh.bFaceCell), Sf=f64(mesh.Sf), Cf=f64(mesh.Cf),
                magSf=f64(mesh.magSf), C=f64(mesh.C), V=f64(mesh.V), w=f64(mesh.weights), dc=f64(mesh.deltaCoeffs),
                ndc=f64(mesh.nonOrthDeltaCoeffs), cv=f64(mesh.corrVec), bdc=f64(mesh.bDeltaCoeffs),
                uType=i32(uType), uValue=f64(uValue), pType=i32(pType), pValue=f64(pValue),
                nuEff=None if nuEff is None else f64(nuEff), nuEffB=None if nuEffB is None else f64(nuEffB))
    a = lambda k: None if keep[k] is None else keep[k].ctypes.data
    cs = SimpleCase(mesh.nCells, mesh.nInternal, mesh.nBFaces, a("own"), a("nei"), a("bfc"), a("Sf"), a("Cf"),
                    a("magSf"), a("C"), a("V"), a("w"), a("dc"), a("ndc"), a("cv"), a("bdc"), a("uType"),
                    a("uValue"), a("pType"), a("pValue"), nu, a("nuEff"), a("nuEffB"), relaxU, relaxP,
                    momentumPredictor, linearUpwind, correctedLaplacian, nNonOrthCorr, tolU, relTolU, maxIterU,
                    minIterU, tolP, relTolP, maxIterP, minIterP, pSolver, None if aggl is None else aggl.h,
                    amgPre, amgPost, amgFlexible, amgSingle, amgScale, pRefCell, pRefValue)
    cs._keep = keep
    cs._aggl = aggl
    return cs


def simple_step(cs, st):
    """one SIMPLE iteration in place on the SimpleState; returns the SimpleResult"""
    f = SimpleFields(*[x.ctypes.data for x in st.U + st.Ub + [st.p, st.pb, st.phi, st.phiB]])
    r = SimpleResult()
    rc = lib().orc_simple_step(ctypes.byref(cs), ctypes.byref(f), ctypes.byref(r))
    if rc != 0:
        raise RuntimeError("oracle SIMPLE step refused (adjustPhi)")
    return r


# ---- the standard k-epsilon model (oracle/kepsilon.c) ----
class KEpsilonCoeffs(ctypes.Structure):
    _fields_ = [(n, _D) for n in ("Cmu", "C1", "C2", "C3", "sigmaK", "sigmaEps", "kappa", "E")]


class KEpsilonCase(ctypes.Structure):
    _fields_ = [("nCells", _I), ("nFaces", _I), ("nBFaces", _I), ("owner", _VP), ("neighbour", _VP),
                ("bFaceCell", _VP), ("Sf", _VP), ("magSf", _VP), ("V", _VP), ("weights", _VP),
                ("deltaCoeffs", _VP), ("bDeltaCoeffs", _VP), ("isWall", _VP), ("yWall", _VP), ("kType", _VP),
                ("epsType", _VP), ("kValue", _VP), ("epsValue", _VP), ("nu", _D), ("relaxEps", _D),
                ("relaxK", _D), ("tol", _D), ("relTol", _D), ("maxIter", _I), ("minIter", _I),
                ("co", KEpsilonCoeffs)]


class KEpsilonResult(ctypes.Structure):
    _fields_ = [("epsInitial", _D), ("epsFinal", _D), ("kInitial", _D), ("kFinal", _D), ("epsIters", _I),
                ("kIters", _I), ("nWallCells", _I)]


def kepsilon_coeffs(**kw):
    co = KEpsilonCoeffs()
    lib().orc_kepsilon_default_coeffs(ctypes.byref(co))
    for k, v in kw.items():
        setattr(co, k, v)
    return co


def near_wall_dist(mesh, isWall):
    """y per boundary face of the wall patches (0 elsewhere)"""
    nB = mesh.nFaces - mesh.nInternal
    y = np.zeros(nB)
    lib().orc_near_wall_dist(mesh.nInternal, nB, p(mesh.points), p(mesh.faceOffsets), p(mesh.faceVerts),
                             p(i32(mesh.bFaceCell)), p(f64(mesh.C)), p(i32(isWall)), p(y))
    return y


def kepsilon_case(mesh, isWall, yWall, kType, kValue, epsType, epsValue, nu=1e-5, relaxEps=1.0, relaxK=1.0,
                  tol=1e-10, relTol=0.0, maxIter=2000, minIter=0, co=None):
    keep = dict(own=mesh.own(), nei=i32(mesh.neighbour), bfc=i32(mesh.bFaceCell), Sf=f64(mesh.Sf),
                magSf=f64(mesh.magSf), V=f64(mesh.V), w=f64(mesh.weights), dc=f64(mesh.deltaCoeffs),
                bdc=f64(mesh.bDeltaCoeffs), isWall=i32(isWall), y=f64(yWall), kT=i32(kType), eT=i32(epsType),
                kV=f64(kValue), eV=f64(epsValue))
    a = lambda k: keep[k].ctypes.data
    cs = KEpsilonCase(mesh.nCells, mesh.nInternal, mesh.nFaces - mesh.nInternal, a("own"), a("nei"), a("bfc"),
                      a("Sf"), a("magSf"), a("V"), a("w"), a("dc"), a("bdc"), a("isWall"), a("y"), a("kT"), a("eT"),
                      a("kV"), a("eV"), nu, relaxEps, relaxK, tol, relTol, maxIter, minIter,
                      co if co is not None else kepsilon_coeffs())
    cs._keep = keep
    return cs


def gbynu(gradU9):
    """gradU component-major [9, n] -> GbyNu [n]"""
    g = f64(gradU9)
    n = g.size // 9
    out = np.zeros(n)
    lib().orc_gbynu(n, p(g), p(out))
    return out


def kepsilon_eqns(cs, k, eps, nut, nutb, G, GbyNu, phi, phiB):
    nC, nI, nB = cs.nCells, cs.nFaces, cs.nBFaces
    out = [np.zeros(n) for n in (nC, nI, nI, nC, nB, nB, nC, nI, nI, nC, nB, nB)]
    lib().orc_kepsilon_eqns(ctypes.byref(cs), p(f64(k)), None, p(f64(eps)), None, p(f64(nut)),
                            None if nutb is None else p(f64(nutb)), p(f64(G)), p(f64(GbyNu)), p(f64(phi)),
                            p(f64(phiB)), *[p(a) for a in out])
    return out[:6], out[6:]


def kepsilon_correct(cs, U, Ub, phi, phiB, k, kb, eps, epsb, nut, nutb):
    """one correct() in place on the float64 arrays k, kb, eps, epsb, nut, nutb"""
    keepU = [f64(u) for u in U] + [f64(u) for u in Ub]
    Ua = (ctypes.c_void_p * 3)(*[a.ctypes.data for a in keepU[:3]])
    Uba = (ctypes.c_void_p * 3)(*[a.ctypes.data for a in keepU[3:]])
    r = KEpsilonResult()
    for a in (k, kb, eps, epsb, nut, nutb):
        assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    lib().orc_kepsilon_correct(ctypes.byref(cs), Ua, Uba, p(f64(phi)), p(f64(phiB)), p(k), p(kb), p(eps), p(epsb),
                               p(nut), p(nutb), ctypes.byref(r))
    return r
