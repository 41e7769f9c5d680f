"""ctypes binding of libmfv.so (the C ABI in include/mfv.h).

The library is the product; this module only loads it and declares the signatures so Python
callers (tests, bench.py) can drive it with device memory they own (torch tensors on cuda:0 are
plain HBM allocations here). It never falls back to anything: if libmfv.so is missing or does not
load, every call raises.
"""
import atexit
import collections
import ctypes
import os
import re
import weakref

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "libmfv.so")
HEADERS = [os.path.join(ROOT, "include", h) for h in ("mfv.h",)]

MFV_OK = 0
MFV_ERR_INVALID, MFV_ERR_HIP, MFV_ERR_UNSUPPORTED, MFV_ERR_ALLOC, MFV_ERR_STATE = 1, 2, 3, 4, 5
_lib = None

P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double


class MfvError(RuntimeError):
    pass


class Ldu(ctypes.Structure):
    _fields_ = [("dDiag", P), ("dUpper", P), ("dLower", P)]


class AmgControls(ctypes.Structure):
    _fields_ = [("agglomeration", I), ("nCellsInCoarsestLevel", I), ("mergeLevels", I), ("fineLayout", I)]


class PcgControls(ctypes.Structure):
    _fields_ = [("tolerance", D), ("relTol", D), ("maxIter", I), ("minIter", I), ("precision", I),
                ("nPreSweeps", I), ("nPostSweeps", I), ("flexible", I), ("smoother", I), ("checkEvery", I),
                ("nVcycles", I), ("scaleCorrection", I), ("captureVcycle", I)]


class BatchSolve(ctypes.Structure):
    _fields_ = [("mesh", P), ("amg", P), ("dDiag", P), ("dB", P), ("dPsi", P), ("controls", PcgControls)]


AGG_BRAE, AGG_FACEAREAPAIR, AGG_PAIR = 0, 1, 2
P_AMG_PCG, P_GAMG_REF = 0, 1
LAYOUT_AUTO, LAYOUT_ROWS = 0, 1
CACHE_LOADED, CACHE_ABSENT, CACHE_OTHER_MESH_OR_BUILD, CACHE_UNREADABLE = range(4)
PROF_PCG_SPMV64, PROF_PCG_UPDATE, PROF_PCG_ZR, PROF_VCYCLE, PROF_AMG_RESIDUAL0, PROF_AMG_CORR_SMOOTH0, \
    PROF_AMG_RESTRICT0, PROF_AMG_CORR_AMUL0 = range(8)
# the phases of one SIMPLE iteration (mfv_simple_step)
PROF_SIMPLE = {8: "UEqn assembly", 9: "U solve (momentum predictor)", 10: "pressure predictor",
               11: "pressure assembly", 12: "Galerkin", 13: "pressure solve", 14: "flux and U correction",
               15: "turbulence"}
PREC_FP64, PREC_FP32, PREC_FP16 = 64, 32, 16


SMOOTHER_GAUSS_SEIDEL, SMOOTHER_SYM_GAUSS_SEIDEL, SMOOTHER_DIC, SMOOTHER_DIC_GAUSS_SEIDEL = range(4)
_SMOOTHER_NAMES = {"GaussSeidel": 0, "symGaussSeidel": 1, "DIC": 2, "DICGaussSeidel": 3}


class GamgControls(ctypes.Structure):
    _fields_ = [("smoother", I), ("tolerance", D), ("relTol", D), ("maxIter", I), ("minIter", I),
                ("nPreSweeps", I), ("preSweepsLevelMultiplier", I), ("maxPreSweeps", I),
                ("nPostSweeps", I), ("postSweepsLevelMultiplier", I), ("maxPostSweeps", I),
                ("nFinestSweeps", I), ("scaleCorrection", I), ("nVcycles", I)]


def gamg_controls(**kw):
    """GAMGSolver's controls with its defaults; `smoother` may be OpenFOAM's name."""
    c = GamgControls()
    lib().mfv_gamg_default_controls(ctypes.byref(c))
    for k, v in kw.items():
        if k == "smoother" and isinstance(v, str):
            v = _SMOOTHER_NAMES[v]
        if not hasattr(c, k):
            raise AttributeError("gamg controls have no field " + k)
        setattr(c, k, v)
    return c


class ServeRequest(ctypes.Structure):
    _fields_ = [("nx", I), ("ny", I), ("nz", I), ("lx", D), ("ly", D), ("lz", D), ("jitter", D),
                ("seed", ctypes.c_ulonglong), ("nSteps", I), ("tolerance", D), ("arrival", D)]


class ServeResult(ctypes.Structure):
    _fields_ = [("arrival", D), ("admitted", D), ("finished", D), ("status", I), ("nCells", I), ("nSteps", I),
                ("totalIterations", I), ("finalResidual", D), ("fingerprint", ctypes.c_ulonglong), ("resident", I),
                ("batched", I), ("built", D)]


class Perf(ctypes.Structure):
    _fields_ = [("initialResidual", D), ("finalResidual", D), ("nIterations", I), ("converged", I)]


class BicgControls(ctypes.Structure):
    _fields_ = [("tolerance", D), ("relTol", D), ("maxIter", I), ("minIter", I), ("preconditioner", I),
                ("checkEvery", I)]


PRECOND_DILU, PRECOND_DIAGONAL = 0, 1
BC_ZERO_GRADIENT, BC_FIXED_VALUE, BC_EMPTY = 0, 1, 2


class Geometry(ctypes.Structure):
    _fields_ = [(n, P) for n in ("dSf", "dCf", "dMagSf", "dC", "dV", "dWeights", "dDeltaCoeffs",
                                 "dNonOrthDeltaCoeffs", "dCorrVec", "dBDeltaCoeffs")]


class SimpleControls(ctypes.Structure):
    _fields_ = [("nu", D), ("relaxU", D), ("relaxP", D), ("momentumPredictor", I), ("linearUpwind", I),
                ("correctedLaplacian", I), ("nNonOrthCorr", I), ("pRefCell", I), ("pRefValue", D),
                ("U", BicgControls), ("p", PcgControls), ("mergeLevels", I), ("pSolver", I)]


class KEpsilonCoeffs(ctypes.Structure):
    _fields_ = [(n, D) for n in ("Cmu", "C1", "C2", "C3", "sigmaK", "sigmaEps", "kappa", "E")]


class KEpsilonControls(ctypes.Structure):
    _fields_ = [("nu", D), ("relaxEps", D), ("relaxK", D), ("solver", BicgControls), ("coeffs", KEpsilonCoeffs)]


class KEpsilonFields(ctypes.Structure):
    _fields_ = [(n, P) for n in ("dUx", "dUy", "dUz", "dUbx", "dUby", "dUbz", "dPhi", "dPhiB", "dK", "dKb", "dEps",
                                 "dEpsb", "dNut", "dNutb")]


class KEpsilonResult(ctypes.Structure):
    _fields_ = [("epsInitial", D), ("epsFinal", D), ("kInitial", D), ("kFinal", D), ("epsIters", I),
                ("kIters", I), ("nWallCells", I)]


class SimpleFields(ctypes.Structure):
    _fields_ = [(n, P) for n in ("dUx", "dUy", "dUz", "dUbx", "dUby", "dUbz", "dP", "dPb", "dPhi", "dPhiB")]


class SimpleResult(ctypes.Structure):
    _fields_ = [("UInitial", D * 3), ("UFinal", D * 3), ("UIters", I * 3), ("pInitial", D), ("pFinal", D),
                ("pIters", I * 8), ("massCorr", D)]


_SIGS = {
    "mfv_ctx_create": [I, P, P],
    "mfv_ctx_destroy": [P],
    "mfv_last_error": [P],
    "mfv_ctx_stream": [P],
    "mfv_ctx_sync": [P],
    "mfv_ctx_set_sequential_reduce": [P, I],
    "mfv_mesh_create": [P, I, I, P, P, I, P, P],
    "mfv_mesh_destroy": [P],
    "mfv_mesh_addressing": [P, P, P, P, P, P],
    "mfv_laplacian_coeffs": [P, P, P, P, P, P, P, P],
    "mfv_laplacian_bcoeffs": [P, P, P, P, P, P, P, P, P],
    "mfv_div_upwind_coeffs": [P, P, P, P, P, P],
    "mfv_div_bcoeffs": [P, P, P, P, P, P, P],
    "mfv_div_central_coeffs": [P, P, P, P, P, P, P],
    "mfv_fold": [P, P, P, P, P, P, P, P],
    "mfv_set_reference": [P, I, P, P, I, D],
    "mfv_face_flux": [P, P, P, P, P],
    "mfv_amul": [P, P, P, P, P],
    "mfv_residual": [P, P, P, P, P, P],
    "mfv_dot": [P, I, P, P, P],
    "mfv_sum_mag": [P, I, P, P],
    "mfv_norm_factor": [P, P, P, P, P, P],
    "mfv_interpolate": [P, P, P, P, P, P],
    "mfv_div_source": [P, P, P, P, P, D, P],
    "mfv_flux_correct": [P, P, P, P, P, P, D, P],
    "mfv_grad_gauss": [P, P, P, P, P, P, P, P, P],
    "mfv_laplacian_corr": [P, P, P, P, P, P, P, P, P],
    "mfv_laplacian_corr_limited": [P, P, P, P, P, P, P, P, P, D, P, P],
    "mfv_mesh_geometry": [P, P, P, P, P] + [P] * 10,
    "mfv_ctx_profile": [P, I],
    "mfv_ctx_profile_read": [P, I, P, P],
    "mfv_amg_build": [P, P, P, P, P],
    "mfv_amg_destroy": [P],
    "mfv_amg_signature": [I, I, P, P, P, P, P],
    "mfv_amg_save": [P, P, ctypes.c_char_p],
    "mfv_amg_build_or_load": [P, P, P, P, ctypes.c_char_p, I, P, P],
    "mfv_amg_clone": [P, P, P],
    "mfv_amg_num_grids": [P, P],
    "mfv_amg_grid_size": [P, I, P, P],
    "mfv_amg_fine_stencil": [P, P, P],
    "mfv_amg_coarse_mode": [P, I, P, P, P],
    "mfv_amg_pcg_solve_batch": [P, I, P, P],
    "mfv_fold_boundary_diag": [P, P, P, P, P, P, P],
    "mfv_reciprocal_v": [P, P, P, P, P],
    "mfv_boundary_diag_term": [P, P, P, P, P, I, P],
    "mfv_matrix_h": [P, P, P, P, P, P, P, P, P, I, P],
    "mfv_matrix_h3": [P, P, P, P, P, P, P, P, P, P, P],
    "mfv_constrain_hbya": [P, P, P, P, P, P],
    "mfv_vector_flux": [P, P, P, P, P, P, P, P],
    "mfv_boundary_flux": [P, P, P, P, P, P, P, P],
    "mfv_adjust_phi": [P, I, P, P, I, P, P, P],
    "mfv_relax_diag": [P, P, P, P, D, P, P, P, P, P, I, P, P],
    "mfv_linear_upwind_corr": [P, P, P, P, P, I, P, P],
    "mfv_scaled_dev2t": [P, I, P, P, P],
    "mfv_tensor_div_source": [P, P, P, P, P, P, P, P, P, P],
    "mfv_div_dev_reff_explicit": [P, P, P, P, P, P, P, P, P, P, P],
    "mfv_cmpt_max_mag3": [P, I, P, P, P, P],
    "mfv_cmpt_min3": [P, I, P, P, P, P],
    "mfv_relax_field": [P, I, P, P, D],
    "mfv_amg_level_addressing": [P, I, P, P, P, P, P],
    "mfv_amg_galerkin": [P, P, P, P, P, I],
    "mfv_amg_grid_values": [P, P, I, P, P],
    "mfv_amg_vcycle": [P, P, P, P, I, I, I],
    "mfv_pcg_default_controls": [P],
    "mfv_amg_pcg_solve": [P, P, P, P, P, P, P, P, P],
    "mfv_dic_precondition": [P, P, P, P, P],
    "mfv_gauss_seidel": [P, P, P, P, P, I],
    "mfv_pcg_dic_solve": [P, P, P, P, P, P, P, P],
    "mfv_sym_gauss_seidel": [P, P, P, P, P, I],
    "mfv_gamg_default_controls": [P],
    "mfv_faceareapair_weights": [I, P, P, P],
    "mfv_gamg_solve": [P, P, P, P, P, P, P, P],
    "mfv_pcg_gamg_solve": [P, P, P, P, P, P, D, D, I, I, P, P],
    "mfv_gamg_ref_solve": [P, P, P, P, P, P, D, D, I, I, I, I, P],
    "mfv_bicg_default_controls": [P],
    "mfv_pbicgstab_solve": [P, P, P, P, P, P, P, P, P],
    "mfv_pbicgstab_solve_noalloc": [P, P, P, P, P, P, P, P, P],
    "mfv_kepsilon_default_controls": [P],
    "mfv_near_wall_dist": [I, I, P, P, P, P, P, P, P],
    "mfv_kepsilon_create": [P, P, P, P, P, P, P, P, P, P, P],
    "mfv_kepsilon_destroy": [P],
    "mfv_kepsilon_nu_eff": [P, P, P, P, P],
    "mfv_kepsilon_correct": [P, P, P],
    "mfv_simple_step_kepsilon": [P, P, P, P, P, P],
    "mfv_simple_default_controls": [P],
    "mfv_simple_create": [P, P, P, P, P, P, P, P, P],
    "mfv_simple_destroy": [P],
    "mfv_simple_step": [P, P, P, P, P],
    "mfv_server_create": [I, I, D, I, P],
    "mfv_server_destroy": [P],
    "mfv_server_set_resident": [P, D],
    "mfv_server_set_batch": [P, I],
    "mfv_server_request_bytes": [P, P],
    "mfv_server_run": [P, P, I, P],
}
_RESTYPES = {"mfv_last_error": ctypes.c_char_p, "mfv_ctx_stream": P, "mfv_server_request_bytes": D}


def solve_batch(ctx, jobs):
    """whole AMG-PCG solves batched into one launch (Amg.solve_batch_job entries); returns their Perf"""
    arr = (BatchSolve * len(jobs))(*jobs)
    perfs = (Perf * len(jobs))()
    ctx.call("mfv_amg_pcg_solve_batch", len(jobs), arr, perfs)
    return list(perfs)


def serve(requests, nSlots=8, hbmBudgetBytes=200e9, precision=32, device=0, residentBytes=0.0, batchCells=0):
    """run a request stream (list of ServeRequest or dicts of its fields) through the serving loop, with
    residentBytes of the budget keeping systems between requests and requests of at most batchCells cells
    solved in batches; returns (status, [ServeResult])"""
    reqs = (ServeRequest * len(requests))()
    for i, r in enumerate(requests):
        if isinstance(r, dict):
            for k, v in r.items():
                setattr(reqs[i], k, v)
        else:
            reqs[i] = r
    res = (ServeResult * len(requests))()
    srv = P()
    rc = lib().mfv_server_create(device, nSlots, float(hbmBudgetBytes), precision, ctypes.byref(srv))
    if rc != MFV_OK:
        raise MfvError("server_create failed (%d)" % rc)
    try:
        if residentBytes:
            rc = lib().mfv_server_set_resident(srv, float(residentBytes))
            if rc != MFV_OK:
                raise MfvError("server_set_resident failed (%d)" % rc)
        if batchCells:
            rc = lib().mfv_server_set_batch(srv, int(batchCells))
            if rc != MFV_OK:
                raise MfvError("server_set_batch failed (%d)" % rc)
        rc = lib().mfv_server_run(srv, reqs, len(requests), res)
    finally:
        lib().mfv_server_destroy(srv)
    return rc, list(res)


def amg_signature(nCells, owner, neighbour, weights, agglomeration=AGG_BRAE, nCellsInCoarsestLevel=0, mergeLevels=1):
    """the hierarchy fingerprint (host only; no GPU needed)"""
    import numpy as np
    own = np.ascontiguousarray(owner, np.int32)
    nei = np.ascontiguousarray(neighbour, np.int32)
    w = np.ascontiguousarray(weights, np.float64)
    sig = ctypes.c_ulonglong()
    c = AmgControls(agglomeration, nCellsInCoarsestLevel, mergeLevels)
    rc = lib().mfv_amg_signature(int(nCells), own.size, ptr(own), ptr(nei), ptr(w), ctypes.byref(c), ctypes.byref(sig))
    if rc != MFV_OK:
        raise MfvError("amg_signature failed (%d)" % rc)
    return sig.value


def declared_symbols():
    """Every mfv_* function declared in the public headers."""
    names = []
    for h in HEADERS:
        txt = open(h).read()
        names += re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(mfv_\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MfvError("libmfv.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        # torch ships its own HIP runtime under the same sonames; loading it first makes the library bind to that
        # one, so a process that creates a context before touching torch still has a single runtime
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, I)
        _lib = L
    return _lib


def declare(name, args, restype=I):
    """Register a signature for an entry point added after this module was written."""
    _SIGS[name] = args
    if _lib is not None:
        fn = getattr(_lib, name)
        fn.argtypes = args
        fn.restype = restype


_KEEP = collections.deque(maxlen=256)


def ptr(t):
    """Device/host address of a torch tensor or numpy array (None -> NULL). The object is kept
    alive for the next calls so a temporary in an argument list outlives the call using it."""
    if t is None:
        return None
    _KEEP.append(t)
    if hasattr(t, "data_ptr"):
        assert t.is_contiguous()
        return t.data_ptr()
    assert t.flags["C_CONTIGUOUS"]
    return t.ctypes.data


_LIVE = weakref.WeakSet()


def _adopt(parent, child):
    """record that child was built on parent (a context's meshes, a mesh's hierarchies and solvers)"""
    parent.__dict__.setdefault("_children", weakref.WeakSet()).add(child)


def _close_children(obj):
    """close what was built on obj before obj itself: the garbage collector finalises the objects of a
    reference cycle in any order, and the library needs a context and a mesh to outlive what was built on
    them"""
    kids = obj.__dict__.get("_children")
    if kids:
        for k in list(kids):
            k.close()


@atexit.register
def _close_all():
    """Release library objects before the HIP runtime tears down at process exit: meshes and
    hierarchies first, contexts last."""
    objs = list(_LIVE)
    for o in objs:
        if not isinstance(o, Context):
            o.close()
    for o in objs:
        if isinstance(o, Context):
            o.close()
    _KEEP.clear()


class Context:
    def __init__(self, device=0, stream=None):
        self.h = P()
        self._check(lib().mfv_ctx_create(device, stream, ctypes.byref(self.h)), "ctx_create")
        _LIVE.add(self)

    def _check(self, rc, what):
        if rc != MFV_OK:
            msg = lib().mfv_last_error(self.h if isinstance(self, Context) and self.h else None)
            raise MfvError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def call(self, name, *args):
        rc = getattr(lib(), name)(self.h, *args)
        self._check(rc, name)

    def sync(self):
        self.call("mfv_ctx_sync")

    def sequential_reduce(self, on=True):
        self.call("mfv_ctx_set_sequential_reduce", 1 if on else 0)

    def profile(self, on=True):
        self.call("mfv_ctx_profile", 1 if on else 0)

    def profile_read(self, tag):
        n, ms = ctypes.c_int(), ctypes.c_double()
        self.call("mfv_ctx_profile_read", tag, ctypes.byref(n), ctypes.byref(ms))
        return n.value, ms.value

    @property
    def stream(self):
        return lib().mfv_ctx_stream(self.h)

    def close(self):
        if self.h:
            _close_children(self)
            lib().mfv_ctx_destroy(self.h)
            self.h = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Mesh:
    """Device LDU addressing of a mesh (host numpy int32 arrays in)."""

    def __init__(self, ctx, nCells, owner, neighbour, bFaceCell):
        import numpy as np
        self.ctx = ctx
        self.nCells = int(nCells)
        self.owner = np.ascontiguousarray(owner, np.int32)
        self.neighbour = np.ascontiguousarray(neighbour, np.int32)
        self.bFaceCell = np.ascontiguousarray(bFaceCell, np.int32)
        self.nFaces = self.owner.size
        self.nBFaces = self.bFaceCell.size
        self.h = P()
        ctx._check(lib().mfv_mesh_create(ctx.h, self.nCells, self.nFaces, ptr(self.owner), ptr(self.neighbour),
                                         self.nBFaces, ptr(self.bFaceCell), ctypes.byref(self.h)), "mesh_create")
        _LIVE.add(self)
        _adopt(ctx, self)

    def close(self):
        if self.h:
            _close_children(self)
            lib().mfv_mesh_destroy(self.h)
            self.h = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ldu(diag, upper, lower=None):
    return Ldu(ptr(diag), ptr(upper), ptr(lower if lower is not None else upper))


class Amg:
    """AMG hierarchy of a mesh (host agglomeration, device levels)."""

    def __init__(self, mesh, weights, agglomeration=AGG_BRAE, nCellsInCoarsestLevel=0, mergeLevels=1,
                 cacheDir=None, writeCache=False, _handle=None, fineLayout=LAYOUT_AUTO):
        """build the hierarchy; with cacheDir, load <cacheDir>/.mfv_amgcache when it is this mesh's
        (self.how tells what the file held: CACHE_*)"""
        import numpy as np
        self.ctx = mesh.ctx
        self.mesh = mesh
        self.how = None
        self.h = P()
        if _handle is not None:
            self.h = _handle
        else:
            w = np.ascontiguousarray(weights, np.float64)
            assert w.size == mesh.nFaces
            c = AmgControls(agglomeration, nCellsInCoarsestLevel, mergeLevels, fineLayout)
            if cacheDir is None:
                self.ctx._check(lib().mfv_amg_build(self.ctx.h, mesh.h, ptr(w), ctypes.byref(c), ctypes.byref(self.h)),
                                "amg_build")
            else:
                how = ctypes.c_int(-1)
                self.ctx._check(lib().mfv_amg_build_or_load(self.ctx.h, mesh.h, ptr(w), ctypes.byref(c),
                                                            str(cacheDir).encode(), int(writeCache),
                                                            ctypes.byref(self.h), ctypes.byref(how)),
                                "amg_build_or_load")
                self.how = how.value
        _LIVE.add(self)
        _adopt(self.mesh, self)
        n = ctypes.c_int()
        lib().mfv_amg_num_grids(self.h, ctypes.byref(n))
        self.nGrids = n.value

    def clone(self):
        h = P()
        self.ctx._check(lib().mfv_amg_clone(self.ctx.h, self.h, ctypes.byref(h)), "amg_clone")
        return Amg(self.mesh, None, _handle=h)

    def save(self, path):
        self.ctx._check(lib().mfv_amg_save(self.ctx.h, self.h, str(path).encode()), "amg_save")

    def grid_size(self, g):
        nc, nf = ctypes.c_int(), ctypes.c_int()
        lib().mfv_amg_grid_size(self.h, g, ctypes.byref(nc), ctypes.byref(nf))
        return nc.value, nf.value

    def fine_stencil(self):
        """the fine grid's stencil offsets when the outer product reads it as diagonals, else []"""
        m = ctypes.c_int()
        offs = (ctypes.c_int * 4)()
        self.ctx._check(lib().mfv_amg_fine_stencil(self.h, ctypes.byref(m), offs), "amg_fine_stencil")
        return list(offs[:m.value])

    def coarse_mode(self, mode=-1):
        """sets how the coarsest grids run (0 per-grid kernels, 1 one workgroup, 2 one workgroup with the
        coarsest grids staged in LDS; -1 leaves it) and returns (the tail's first grid, the first grid
        staged, the LDS bytes the staged grids take)"""
        t, g, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self.ctx._check(lib().mfv_amg_coarse_mode(self.h, mode, ctypes.byref(t), ctypes.byref(g), ctypes.byref(b)),
                        "amg_coarse_mode")
        return t.value, g.value, b.value

    def level(self, i):
        import numpy as np
        nF, fF = self.grid_size(i)
        nC, cF = self.grid_size(i + 1)
        m, co, cn = np.zeros(nF, np.int32), np.zeros(cF, np.int32), np.zeros(cF, np.int32)
        fr, ff = np.zeros(fF, np.int32), np.zeros(fF, np.int8)
        rc = lib().mfv_amg_level_addressing(self.h, i, ptr(m), ptr(co), ptr(cn), ptr(fr), ptr(ff))
        assert rc == 0
        return dict(map=m, cOwn=co, cNei=cn, faceRestrict=fr, faceFlip=ff, nCoarse=nC, nCoarseFaces=cF)

    def galerkin(self, diag, upper, precision=PREC_FP32):
        self.ctx.call("mfv_amg_galerkin", self.h, ptr(diag), ptr(upper), None, precision)

    def grid_values(self, g):
        import numpy as np
        n, f = self.grid_size(g)
        d, u = np.zeros(n), np.zeros(f)
        self.ctx.call("mfv_amg_grid_values", self.h, g, ptr(d), ptr(u))
        return d, u

    def vcycle(self, r, z, nPre=1, nPost=1, scale=False):
        self.ctx.call("mfv_amg_vcycle", self.h, ptr(r), ptr(z), nPre, nPost, 1 if scale else 0)

    def solve(self, diag, upper, b, psi, normFactor=None, **kw):
        c = PcgControls()
        lib().mfv_pcg_default_controls(ctypes.byref(c))
        for k, v in kw.items():
            setattr(c, k, v)
        A = ldu(diag, upper)
        perf = Perf()
        self.ctx.call("mfv_amg_pcg_solve", self.mesh.h, self.h, ctypes.byref(A), ptr(b), ptr(psi),
                      ptr(normFactor), ctypes.byref(c), ctypes.byref(perf))
        return perf

    def solve_batch_job(self, diag, b, psi, **kw):
        """one entry of a batched solve (solve_batch) on this hierarchy"""
        c = PcgControls()
        lib().mfv_pcg_default_controls(ctypes.byref(c))
        for k, v in kw.items():
            setattr(c, k, v)
        return BatchSolve(self.mesh.h, self.h, ptr(diag), ptr(b), ptr(psi), c)

    def gamg_solve(self, diag, upper, b, psi, **kw):
        """OpenFOAM's GAMGSolver on the folded system (psi in place)."""
        c = gamg_controls(**kw)
        perf = Perf()
        A = ldu(diag, upper)
        self.ctx.call("mfv_gamg_solve", self.mesh.h, self.h, ctypes.byref(A), ptr(b), ptr(psi), ctypes.byref(c),
                      ctypes.byref(perf))
        return perf

    def pcg_gamg_solve(self, diag, upper, b, psi, tol, relTol=0.0, maxIter=1000, minIter=0, **kw):
        """PCG with the GAMG preconditioner: tol..minIter are the PCG's, kw the preconditioner's GAMG
        controls (its tolerance and relTol are the coarsest solve's)."""
        c = gamg_controls(**kw)
        perf = Perf()
        A = ldu(diag, upper)
        self.ctx.call("mfv_pcg_gamg_solve", self.mesh.h, self.h, ctypes.byref(A), ptr(b), ptr(psi), tol, relTol,
                      maxIter, minIter, ctypes.byref(c), ctypes.byref(perf))
        return perf

    def close(self):
        if self.h:
            lib().mfv_amg_destroy(self.h)
            self.h = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ptr3(ts):
    """a host array of three device pointers (per-component arguments)"""
    a = (ctypes.c_void_p * 3)(*[ptr(t) for t in ts])
    _KEEP.append(a)
    return a



def bicg_controls(**kw):
    """PBiCGStab controls with the library's defaults (1e-10, DILU), overridden by keyword"""
    c = BicgControls()
    lib().mfv_bicg_default_controls(ctypes.byref(c))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise AttributeError("bicg controls have no field " + k)
        setattr(c, k, v)
    return c


def simple_controls(U=None, p=None, **kw):
    """one SIMPLE iteration's controls: the library's defaults, then U/p solver fields (dicts) and the rest"""
    c = SimpleControls()
    lib().mfv_simple_default_controls(ctypes.byref(c))
    for k, v in (U or {}).items():
        setattr(c.U, k, v)
    for k, v in (p or {}).items():
        setattr(c.p, k, v)
    for k, v in kw.items():
        if not hasattr(c, k):
            raise AttributeError("simple controls have no field " + k)
        setattr(c, k, v)
    return c


class Simple:
    """mfv_simple: one SIMPLE iteration's workspace and pressure hierarchy on a mesh"""

    def __init__(self, mesh, geometry, uType, uValue, pType, pValue, controls):
        import numpy as np
        self.mesh = mesh
        self._keep = [np.ascontiguousarray(uType, np.int32), np.ascontiguousarray(uValue, np.float64),
                      np.ascontiguousarray(pType, np.int32), np.ascontiguousarray(pValue, np.float64)]
        self.geometry = geometry
        self.controls = controls
        h = P()
        mesh.ctx.call("mfv_simple_create", mesh.h, ctypes.byref(geometry), *[k.ctypes.data for k in self._keep],
                      ctypes.byref(controls), ctypes.byref(h))
        self.h = h.value
        _LIVE.add(self)
        _adopt(self.mesh, self)

    def step(self, U, Ub, p, pb, phi, phiB, nuEff=None, nuEffB=None):
        f = SimpleFields(*[ptr(t) for t in list(U) + list(Ub) + [p, pb, phi, phiB]])
        r = SimpleResult()
        rc = lib().mfv_simple_step(self.h, ctypes.byref(f), ptr(nuEff), ptr(nuEffB), ctypes.byref(r))
        self.mesh.ctx._check(rc, "mfv_simple_step")
        return r

    def close(self):
        if getattr(self, "h", None):
            lib().mfv_simple_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def kepsilon_controls(solver=None, coeffs=None, **kw):
    """k-epsilon controls with the library's defaults, then solver/coeffs fields (dicts) and the rest"""
    c = KEpsilonControls()
    lib().mfv_kepsilon_default_controls(ctypes.byref(c))
    for k, v in (solver or {}).items():
        setattr(c.solver, k, v)
    for k, v in (coeffs or {}).items():
        setattr(c.coeffs, k, v)
    for k, v in kw.items():
        if not hasattr(c, k):
            raise AttributeError("kepsilon controls have no field " + k)
        setattr(c, k, v)
    return c


def near_wall_dist(nInternal, points, faceOffsets, faceVerts, bFaceCell, C, isWall):
    """the library's host nearWallDist: y per boundary face (numpy in, numpy out)"""
    import numpy as np
    pts = np.ascontiguousarray(points, np.float64)
    fo = np.ascontiguousarray(faceOffsets, np.int32)
    fv = np.ascontiguousarray(faceVerts, np.int32)
    bfc = np.ascontiguousarray(bFaceCell, np.int32)
    cc = np.ascontiguousarray(C, np.float64)
    w = np.ascontiguousarray(isWall, np.int32)
    y = np.zeros(bfc.size)
    rc = lib().mfv_near_wall_dist(nInternal, bfc.size, pts.ctypes.data, fo.ctypes.data, fv.ctypes.data,
                                  bfc.ctypes.data, cc.ctypes.data, w.ctypes.data, y.ctypes.data)
    if rc != MFV_OK:
        raise MfvError("mfv_near_wall_dist failed (%d)" % rc)
    return y


class KEpsilon:
    """mfv_kepsilon: the standard k-epsilon model's workspace on a mesh"""

    def __init__(self, mesh, geometry, isWall, yWall, kType, kValue, epsType, epsValue, controls):
        import numpy as np
        self.mesh = mesh
        self._keep = [np.ascontiguousarray(isWall, np.int32), np.ascontiguousarray(yWall, np.float64),
                      np.ascontiguousarray(kType, np.int32), np.ascontiguousarray(kValue, np.float64),
                      np.ascontiguousarray(epsType, np.int32), np.ascontiguousarray(epsValue, np.float64)]
        self.geometry = geometry
        self.controls = controls
        h = P()
        mesh.ctx.call("mfv_kepsilon_create", mesh.h, ctypes.byref(geometry), *[k.ctypes.data for k in self._keep],
                      ctypes.byref(controls), ctypes.byref(h))
        self.h = h.value
        _LIVE.add(self)
        _adopt(self.mesh, self)

    @staticmethod
    def fields(U, Ub, phi, phiB, k, kb, eps, epsb, nut, nutb):
        return KEpsilonFields(*[ptr(t) for t in list(U) + list(Ub) + [phi, phiB, k, kb, eps, epsb, nut, nutb]])

    def correct(self, U, Ub, phi, phiB, k, kb, eps, epsb, nut, nutb):
        f = self.fields(U, Ub, phi, phiB, k, kb, eps, epsb, nut, nutb)
        r = KEpsilonResult()
        self.mesh.ctx._check(lib().mfv_kepsilon_correct(self.h, ctypes.byref(f), ctypes.byref(r)),
                             "mfv_kepsilon_correct")
        return r

    def close(self):
        if getattr(self, "h", None):
            lib().mfv_kepsilon_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def simple_step_kepsilon(sim, ke, U, Ub, p, pb, phi, phiB, k, kb, eps, epsb, nut, nutb):
    """one turbulent SIMPLE iteration (nuEff from nut, the step, the model's correct())"""
    f = SimpleFields(*[ptr(t) for t in list(U) + list(Ub) + [p, pb, phi, phiB]])
    kf = KEpsilon.fields(U, Ub, phi, phiB, k, kb, eps, epsb, nut, nutb)
    r, kr = SimpleResult(), KEpsilonResult()
    rc = lib().mfv_simple_step_kepsilon(sim.h, ctypes.byref(f), ke.h, ctypes.byref(kf), ctypes.byref(r),
                                        ctypes.byref(kr))
    sim.mesh.ctx._check(rc, "mfv_simple_step_kepsilon")
    return r, kr
