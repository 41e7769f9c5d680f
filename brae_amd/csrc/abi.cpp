// The C ABI (include/mfv.h): argument checks, exception-to-status translation, the context and
// the mesh addressing. Numerical work is in the kernels_*.hip launchers and the solver drivers.
#include "amg.hpp"
#include "mfv.h"
#include "mfv_internal.hpp"
#include "simple.hpp"

#include <atomic>
#include <cmath>
#include <cstring>
#include <new>

// A context and a mesh are reference counted: the caller's handle holds one reference and every object
// built on them (a mesh on its context; a hierarchy, SIMPLE or kEpsilon workspace on its mesh and context)
// another, so the destroy calls may come in any order and the last one tears down.
struct mfv_ctx_s
{
    mfv::Ctx c;
    std::atomic<int> refs{1};
};

struct mfv_mesh_s
{
    mfv_ctx ctx;
    mfv::Mesh m;
    std::atomic<int> refs{1};
};

struct mfv_amg_s
{
    mfv_ctx ctx;
    mfv_mesh mesh;
    mfv::Amg a;
};

struct mfv_simple_s
{
    mfv_ctx ctx;
    mfv_mesh mesh;
    std::unique_ptr<mfv::SimpleSolver> s;
};

struct mfv_kepsilon_s
{
    mfv_ctx ctx;
    mfv_mesh mesh;
    std::unique_ptr<mfv::KEpsilon> k;
    mfv::DBuf<double> nuEffC, nuEffB;   // the turbulent step's nu + nut
};

namespace mfv {

namespace {

thread_local std::string tlsError;

template <typename F>
int guarded(mfv_ctx ctx, F&& f)
{
    try
    {
        f();
        if (ctx) ctx->c.lastError.clear();
        return MFV_OK;
    }
    catch (const Error& e)
    {
        if (ctx) ctx->c.lastError = e.what();
        tlsError = e.what();
        return e.code;
    }
    catch (const std::bad_alloc&)
    {
        if (ctx) ctx->c.lastError = "mfv: host allocation failed";
        tlsError = "mfv: host allocation failed";
        return MFV_ERR_ALLOC;
    }
    catch (const std::exception& e)
    {
        if (ctx) ctx->c.lastError = e.what();
        tlsError = e.what();
        return MFV_ERR_INVALID;
    }
}

void need(bool ok, const char* what)
{
    if (!ok) invalid(what);
}

} // namespace

// ownerStart, stable losort by neighbour, boundary gather (device_ldu.cuh:109-146)
void buildMesh(Ctx& c, Mesh& m, int nCells, int nFaces, const int* owner, const int* neighbour, int nB,
               const int* bFaceCell)
{
    if (nCells <= 0) invalid("mesh: nCells must be positive");
    if (nFaces < 0 || nB < 0) invalid("mesh: negative face count");
    if (nFaces && (!owner || !neighbour)) invalid("mesh: owner/neighbour missing");
    if (nB && !bFaceCell) invalid("mesh: bFaceCell missing");
    m.nCells = nCells;
    m.nFaces = nFaces;
    m.nBFaces = nB;
    m.band = 0;
    m.hOwner.assign(owner, owner + nFaces);
    m.hNeighbour.assign(neighbour, neighbour + nFaces);
    m.hBFaceCell.assign(bFaceCell, bFaceCell + nB);
    std::vector<label> os(nCells + 1, 0), ls(nCells + 1, 0), lo(nFaces), loOwn(nFaces);
    for (int f = 0; f < nFaces; ++f)
    {
        const int o = owner[f], n = neighbour[f];
        if (o < 0 || n >= nCells || o >= n)
            invalid("mesh: internal face " + std::to_string(f) + " is not owner < neighbour within [0, nCells)");
        if (f > 0 && o < owner[f - 1])
            invalid("mesh: internal face " + std::to_string(f) +
                    " breaks OpenFOAM's upper-triangular order (faces sorted by owner)");
        ++os[o + 1];
        ++ls[n + 1];
        if (n - o > m.band) m.band = n - o;
    }
    for (int i = 0; i < nCells; ++i)
    {
        os[i + 1] += os[i];
        ls[i + 1] += ls[i];
    }
    {
        std::vector<label> at(ls.begin(), ls.end() - 1);
        for (int f = 0; f < nFaces; ++f) lo[at[neighbour[f]]++] = f;
    }
    for (int k = 0; k < nFaces; ++k) loOwn[k] = owner[lo[k]];
    std::vector<label> bs(nCells + 1, 0), bf(nB);
    for (int i = 0; i < nB; ++i)
    {
        if (bFaceCell[i] < 0 || bFaceCell[i] >= nCells) invalid("mesh: boundary face cell out of range");
        ++bs[bFaceCell[i] + 1];
    }
    for (int i = 0; i < nCells; ++i) bs[i + 1] += bs[i];
    {
        std::vector<label> at(bs.begin(), bs.end() - 1);
        for (int i = 0; i < nB; ++i) bf[at[bFaceCell[i]]++] = i;
    }
    m.owner.upload(owner, nFaces, c.stream);
    m.neighbour.upload(neighbour, nFaces, c.stream);
    m.ownerStart.upload(os, c.stream);
    m.losortStart.upload(ls, c.stream);
    m.losort.upload(lo, c.stream);
    m.losortOwner.upload(loOwn, c.stream);
    m.bFaceCell.upload(bFaceCell, nB, c.stream);
    m.bndStart.upload(bs, c.stream);
    m.bndFaces.upload(bf, c.stream);
    hipCheck(hipStreamSynchronize(c.stream), "mesh upload");
}

} // namespace mfv

using namespace mfv;

namespace {

LduView view(const mfv_mesh mesh, const mfv_ldu* A)
{
    need(A && A->dDiag && (A->dUpper || mesh->m.nFaces == 0), "matrix: diag/upper missing");
    return LduView{mesh->m.nCells, mesh->m.nFaces, A->dDiag, A->dUpper, A->dLower ? A->dLower : A->dUpper};
}

void needMesh(mfv_ctx ctx, mfv_mesh mesh)
{
    need(ctx != nullptr, "null context");
    need(mesh != nullptr, "null mesh");
    need(mesh->ctx == ctx, "mesh belongs to another context");
}

void putPerf(const PcgPerf& r, mfv_perf* perf)
{
    if (!perf) return;
    perf->initialResidual = r.initialResidual;
    perf->finalResidual = r.finalResidual;
    perf->nIterations = r.nIterations;
    perf->converged = r.converged ? 1 : 0;
}

GamgControls gamgControls(const mfv_gamg_controls& c)
{
    if (c.smoother < 0 || c.smoother > 3)
        throw Error(kErrUnsupported, "mfv: gamg: smoother " + std::to_string(c.smoother) +
                                         " is not carried (GaussSeidel, symGaussSeidel, DIC, DICGaussSeidel)");
    GamgControls g;
    g.smoother = (OfSmoother)c.smoother;
    g.tolerance = c.tolerance;
    g.relTol = c.relTol;
    g.maxIter = c.maxIter;
    g.minIter = c.minIter;
    g.nPreSweeps = c.nPreSweeps;
    g.preSweepsLevelMultiplier = c.preSweepsLevelMultiplier;
    g.maxPreSweeps = c.maxPreSweeps;
    g.nPostSweeps = c.nPostSweeps;
    g.postSweepsLevelMultiplier = c.postSweepsLevelMultiplier;
    g.maxPostSweeps = c.maxPostSweeps;
    g.nFinestSweeps = c.nFinestSweeps;
    g.scaleCorrection = c.scaleCorrection != 0;
    g.nVcycles = c.nVcycles;
    return g;
}

LduView gamgArgs(mfv_ctx ctx, mfv_mesh mesh, mfv_amg amg, const mfv_ldu* A, const double* dB, const double* dPsi,
                 const mfv_gamg_controls* controls)
{
    needMesh(ctx, mesh);
    need(amg && amg->ctx == ctx && amg->mesh == mesh, "gamg: hierarchy of another mesh or context");
    need(dB && dPsi && controls && dB != dPsi, "gamg: null or aliased argument");
    const LduView v = view(mesh, A);
    // the reference refuses an asymmetric matrix (the smoothers and the coarsest solver differ)
    if (v.lower != v.upper) throw Error(kErrUnsupported, "mfv: gamg: an asymmetric matrix is not carried");
    return v;
}

} // namespace

namespace {
PcgControls pcgControls(const mfv_pcg_controls& c);   // the C controls checked and converted
}

extern "C" {

int mfv_ctx_create(int device, void* hStream, mfv_ctx* out)
{
    if (!out) return MFV_ERR_INVALID;
    *out = nullptr;
    mfv_ctx ctx = new (std::nothrow) mfv_ctx_s();
    if (!ctx) return MFV_ERR_ALLOC;
    const int rc = guarded(nullptr, [&] {
        hipCheck(hipSetDevice(device), "set device");
        ctx->c.device = device;
        staticBicgScratchInit(device);   // the no-allocation solve's workspace address on this device
        if (hStream)
            ctx->c.stream = static_cast<hipStream_t>(hStream);
        else
        {
            hipCheck(hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking), "stream create");
            ctx->c.ownsStream = true;
        }
        ctx->c.partials.resize((size_t)8 * kMaxReduceBlocks);
        ctx->c.scalars.resize(64);
        hipCheck(hipMemsetAsync(ctx->c.scalars.p, 0, 64 * sizeof(double), ctx->c.stream), "scalar clear");
        hipCheck(hipHostMalloc((void**)&ctx->c.mailbox, 64 * sizeof(double), hipHostMallocDefault), "mailbox");
        int cu = 0;
        hipCheck(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device), "cu count");
        ctx->c.cuCount = cu > 0 ? cu : 256;
        hipCheck(hipStreamSynchronize(ctx->c.stream), "ctx sync");
    });
    if (rc != MFV_OK)
    {
        if (ctx->c.mailbox) (void)hipHostFree(ctx->c.mailbox);
        if (ctx->c.ownsStream && ctx->c.stream) (void)hipStreamDestroy(ctx->c.stream);
        delete ctx;
        return rc;
    }
    *out = ctx;
    return MFV_OK;
}

extern "C++" {
namespace {

void retainCtx(mfv_ctx ctx) { ctx->refs.fetch_add(1); }

void releaseCtx(mfv_ctx ctx)
{
    if (ctx->refs.fetch_sub(1) != 1) return;
    (void)hipStreamSynchronize(ctx->c.stream);
    for (auto& v : ctx->c.prof)
        for (auto& e : v)
        {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
    for (hipEvent_t e : ctx->c.eventPool) (void)hipEventDestroy(e);
    if (ctx->c.mailbox) (void)hipHostFree(ctx->c.mailbox);
    if (ctx->c.ownsStream) (void)hipStreamDestroy(ctx->c.stream);
    delete ctx;
}

void retainMesh(mfv_mesh mesh) { mesh->refs.fetch_add(1); }

void releaseMesh(mfv_mesh mesh)
{
    if (mesh->refs.fetch_sub(1) != 1) return;
    const mfv_ctx ctx = mesh->ctx;
    (void)hipStreamSynchronize(ctx->c.stream);
    delete mesh;
    releaseCtx(ctx);
}

} // namespace
} // extern "C++"

int mfv_ctx_destroy(mfv_ctx ctx)
{
    if (ctx) releaseCtx(ctx);
    return MFV_OK;
}

const char* mfv_last_error(mfv_ctx ctx)
{
    return ctx ? ctx->c.lastError.c_str() : tlsError.c_str();
}

void* mfv_ctx_stream(mfv_ctx ctx)
{
    return ctx ? (void*)ctx->c.stream : nullptr;
}

int mfv_ctx_sync(mfv_ctx ctx)
{
    return guarded(ctx, [&] {
        need(ctx != nullptr, "null context");
        hipCheck(hipStreamSynchronize(ctx->c.stream), "sync");
    });
}

int mfv_ctx_set_sequential_reduce(mfv_ctx ctx, int sequential)
{
    return guarded(ctx, [&] {
        need(ctx != nullptr, "null context");
        ctx->c.sequentialReduce = sequential != 0;
    });
}

int mfv_mesh_create(mfv_ctx ctx, int nCells, int nFaces, const int* hOwner, const int* hNeighbour, int nBFaces,
                    const int* hBFaceCell, mfv_mesh* out)
{
    if (!out) return MFV_ERR_INVALID;
    *out = nullptr;
    mfv_mesh mesh = nullptr;
    const int rc = guarded(ctx, [&] {
        need(ctx != nullptr, "null context");
        mesh = new mfv_mesh_s();
        mesh->ctx = ctx;
        buildMesh(ctx->c, mesh->m, nCells, nFaces, hOwner, hNeighbour, nBFaces, hBFaceCell);
    });
    if (rc != MFV_OK)
    {
        delete mesh;
        return rc;
    }
    retainCtx(ctx);
    *out = mesh;
    return MFV_OK;
}

int mfv_mesh_destroy(mfv_mesh mesh)
{
    if (mesh) releaseMesh(mesh);
    return MFV_OK;
}

int mfv_mesh_addressing(mfv_mesh mesh, const int** dOwner, const int** dNeighbour, const int** dOwnerStart,
                        const int** dLosortStart, const int** dLosort)
{
    if (!mesh) return MFV_ERR_INVALID;
    if (dOwner) *dOwner = mesh->m.owner.p;
    if (dNeighbour) *dNeighbour = mesh->m.neighbour.p;
    if (dOwnerStart) *dOwnerStart = mesh->m.ownerStart.p;
    if (dLosortStart) *dLosortStart = mesh->m.losortStart.p;
    if (dLosort) *dLosort = mesh->m.losort.p;
    return MFV_OK;
}

int mfv_laplacian_coeffs(mfv_ctx ctx, mfv_mesh mesh, const double* dDeltaCoeffs, const double* dGammaf,
                         const double* dMagSf, double* dDiag, double* dUpper, double* dLower)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(dDeltaCoeffs && dGammaf && dMagSf && dDiag && dUpper, "laplacian: null argument");
        launchLaplacian(ctx->c, mesh->m, dDeltaCoeffs, dGammaf, dMagSf, dDiag, dUpper, dLower ? dLower : dUpper);
    });
}

int mfv_laplacian_bcoeffs(mfv_ctx ctx, mfv_mesh mesh, const int* dBcType, const double* dBcValue,
                          const double* dGammab, const double* dBMagSf, const double* dBDeltaCoeffs, double* dIC,
                          double* dBC)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(!mesh->m.nBFaces || (dBcType && dBcValue && dGammab && dBMagSf && dBDeltaCoeffs && dIC && dBC),
             "laplacian boundary coefficients: null argument");
        launchLaplacianBCoeffs(ctx->c, mesh->m.nBFaces, dBcType, dBcValue, dGammab, dBMagSf, dBDeltaCoeffs, dIC, dBC);
    });
}

int mfv_div_upwind_coeffs(mfv_ctx ctx, mfv_mesh mesh, const double* dPhi, double* dDiag, double* dUpper,
                          double* dLower)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(dPhi && dDiag && dUpper && dLower, "div: null argument");
        need(dUpper != dLower, "div: an upwind matrix is asymmetric; upper and lower must be separate buffers");
        launchDivUpwind(ctx->c, mesh->m, dPhi, dDiag, dUpper, dLower);
    });
}

int mfv_div_central_coeffs(mfv_ctx ctx, mfv_mesh mesh, const double* dPhi, const double* dWeights, double* dDiag,
                           double* dUpper, double* dLower)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(dPhi && dWeights && dDiag && dUpper && dLower, "div: null argument");
        need(dUpper != dLower, "div: a convection matrix is asymmetric; upper and lower must be separate buffers");
        launchDivCentral(ctx->c, mesh->m, dPhi, dWeights, dDiag, dUpper, dLower);
    });
}

int mfv_div_bcoeffs(mfv_ctx ctx, mfv_mesh mesh, const int* dBcType, const double* dBcValue, const double* dPhib,
                    double* dIC, double* dBC)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(!mesh->m.nBFaces || (dBcType && dBcValue && dPhib && dIC && dBC), "div boundary: null argument");
        launchDivBCoeffs(ctx->c, mesh->m.nBFaces, dBcType, dBcValue, dPhib, dIC, dBC);
    });
}

int mfv_fold(mfv_ctx ctx, mfv_mesh mesh, const double* dDiag, const double* dSource, const double* dIC,
             const double* dBC, double* dDiagC, double* dB)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(dDiag && dSource && dDiagC && dB, "fold: null argument");
        need(!mesh->m.nBFaces || (dIC && dBC), "fold: boundary coefficients missing");
        launchFold(ctx->c, mesh->m, dDiag, dSource, dIC, dBC, dDiagC, dB);
    });
}

int mfv_set_reference(mfv_ctx ctx, int nCells, double* dDiag, double* dB, int refCell, double refValue)
{
    return guarded(ctx, [&] {
        need(ctx != nullptr, "null context");
        need(nCells >= 0, "setReference: negative nCells");
        if (refCell < 0) return;
        need(dDiag && dB, "setReference: null argument");
        if (refCell >= nCells)
            invalid("setReference: refCell " + std::to_string(refCell) + " is outside the " + std::to_string(nCells) +
                    " cells");
        launchSetReference(ctx->c, dDiag, dB, refCell, refValue);
    });
}

int mfv_face_flux(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* dPsi, double* dFlux)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(dPsi && dFlux, "face flux: null argument");
        launchFaceFlux(ctx->c, mesh->m, view(mesh, A), dPsi, dFlux);
    });
}

int mfv_amul(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* dPsi, double* dApsi)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(dPsi && dApsi, "amul: null argument");
        launchAmul(ctx->c, mesh->m, view(mesh, A), dPsi, dApsi);
    });
}

int mfv_residual(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* dPsi, const double* dB, double* dR)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(dPsi && dB && dR, "residual: null argument");
        launchResidual(ctx->c, mesh->m, view(mesh, A), dPsi, dB, dR);
    });
}

int mfv_dot(mfv_ctx ctx, int n, const double* dA, const double* dB, double* dResult)
{
    return guarded(ctx, [&] {
        need(ctx != nullptr, "null context");
        need(n >= 0 && (n == 0 || (dA && dB)) && dResult, "dot: bad argument");
        launchDot(ctx->c, n, dA, dB, dResult, 0);
    });
}

int mfv_sum_mag(mfv_ctx ctx, int n, const double* dA, double* dResult)
{
    return guarded(ctx, [&] {
        need(ctx != nullptr, "null context");
        need(n >= 0 && (n == 0 || dA) && dResult, "sum_mag: bad argument");
        launchSumMag(ctx->c, n, dA, dResult, 0);
    });
}

int mfv_norm_factor(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* dPsi, const double* dB,
                    double* dNormFactor)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(dPsi && dB && dNormFactor, "normFactor: null argument");
        launchNormFactor(ctx->c, mesh->m, view(mesh, A), dPsi, dB, dNormFactor);
    });
}


extern "C++" {
namespace {

AmgBuildControls buildControls(const mfv_amg_controls* controls)
{
    AmgBuildControls bc;
    if (!controls) return bc;
    if (controls->agglomeration != MFV_AGG_BRAE && controls->agglomeration != MFV_AGG_FACEAREAPAIR &&
        controls->agglomeration != MFV_AGG_PAIR)
        throw Error(kErrUnsupported, "mfv: amg build: unknown agglomeration " + std::to_string(controls->agglomeration));
    if (controls->agglomeration == MFV_AGG_PAIR && controls->mergeLevels > 1)
        throw Error(kErrUnsupported, "mfv: amg build: gamg.cu's levels are single pairwise passes (mergeLevels 1)");
    if (controls->mergeLevels > 4) throw Error(kErrUnsupported, "mfv: amg build: mergeLevels above 4 is not carried");
    bc.mergeLevels = std::max(1, controls->mergeLevels);
    if (controls->fineLayout != MFV_LAYOUT_AUTO && controls->fineLayout != MFV_LAYOUT_ROWS)
        throw Error(kErrUnsupported, "mfv: amg build: unknown fine layout " + std::to_string(controls->fineLayout));
    bc.fineRows = controls->fineLayout == MFV_LAYOUT_ROWS;
    bc.kind = controls->agglomeration == MFV_AGG_BRAE   ? AggKind::Brae
              : controls->agglomeration == MFV_AGG_PAIR ? AggKind::Pair
                                                        : AggKind::OpenFOAM;
    bc.target = controls->nCellsInCoarsestLevel > 0 ? controls->nCellsInCoarsestLevel
                                                    : (bc.kind == AggKind::Brae ? 64 : 10);
    return bc;
}

// runs `fill` on a new hierarchy of `mesh`; hands it out only when fill succeeded
template <typename F>
int newAmg(mfv_ctx ctx, mfv_mesh mesh, mfv_amg* out, F fill)
{
    if (!out) return MFV_ERR_INVALID;
    *out = nullptr;
    mfv_amg amg = nullptr;
    const int rc = guarded(ctx, [&] {
        needMesh(ctx, mesh);
        amg = new mfv_amg_s();
        amg->ctx = ctx;
        amg->mesh = mesh;
        fill(amg);
    });
    if (rc != MFV_OK)
    {
        delete amg;
        return rc;
    }
    retainCtx(ctx);
    retainMesh(mesh);
    *out = amg;
    return MFV_OK;
}

} // namespace
} // extern "C++"

int mfv_amg_build(mfv_ctx ctx, mfv_mesh mesh, const double* hFaceWeights, const mfv_amg_controls* controls,
                  mfv_amg* out)
{
    return newAmg(ctx, mesh, out, [&](mfv_amg amg) {
        need(hFaceWeights != nullptr || mesh->m.nFaces == 0, "amg build: face weights missing");
        const AmgBuildControls bc = buildControls(controls);
        std::vector<double> w(hFaceWeights, hFaceWeights + mesh->m.nFaces);
        amgBuild(ctx->c, mesh->m, amg->a, w, bc);
        amg->a.signature = amgSignature(mesh->m.hOwner, mesh->m.hNeighbour, w, mesh->m.nCells, bc);
    });
}

int mfv_amg_signature(int nCells, int nFaces, const int* hOwner, const int* hNeighbour, const double* hFaceWeights,
                      const mfv_amg_controls* controls, unsigned long long* signature)
{
    if (!signature || nCells <= 0 || nFaces < 0 || (nFaces && (!hOwner || !hNeighbour || !hFaceWeights)))
        return MFV_ERR_INVALID;
    try
    {
        const AmgBuildControls bc = buildControls(controls);
        *signature = amgSignature(std::vector<int>(hOwner, hOwner + nFaces),
                                  std::vector<int>(hNeighbour, hNeighbour + nFaces),
                                  std::vector<double>(hFaceWeights, hFaceWeights + nFaces), nCells, bc);
        return MFV_OK;
    }
    catch (const Error& e)
    {
        return e.code;
    }
    catch (...)
    {
        return MFV_ERR_INVALID;
    }
}

int mfv_amg_save(mfv_ctx ctx, mfv_amg amg, const char* path)
{
    return guarded(ctx, [&] {
        need(amg && amg->ctx == ctx, "amg save: hierarchy of another context");
        need(path && *path, "amg save: no path");
        amgSave(amg->a, amg->a.signature, path);
    });
}

int mfv_amg_build_or_load(mfv_ctx ctx, mfv_mesh mesh, const double* hFaceWeights, const mfv_amg_controls* controls,
                          const char* cacheDir, int writeCache, mfv_amg* out, int* how)
{
    return newAmg(ctx, mesh, out, [&](mfv_amg amg) {
        need(hFaceWeights != nullptr || mesh->m.nFaces == 0, "amg build: face weights missing");
        need(cacheDir && *cacheDir, "amg build_or_load: no cache directory");
        const AmgBuildControls bc = buildControls(controls);
        std::vector<double> w(hFaceWeights, hFaceWeights + mesh->m.nFaces);
        const uint64_t sig = amgSignature(mesh->m.hOwner, mesh->m.hNeighbour, w, mesh->m.nCells, bc);
        const std::string path = std::string(cacheDir) + "/.mfv_amgcache";
        std::vector<AggLevel> levels;
        AmgBuildControls stored;
        CacheRead r = amgReadLevels(path, sig, levels, stored);
        stored.fineRows = bc.fineRows;
        if (r == CacheRead::Loaded)
        {
            try
            {
                amgBuildFrom(ctx->c, mesh->m, amg->a, std::move(levels), stored);
            }
            catch (const Error& e)
            {
                if (e.code != kErrInvalid) throw;
                r = CacheRead::Unreadable;   // a file whose levels do not fit this mesh: build afresh
            }
        }
        if (r != CacheRead::Loaded)
        {
            amgBuild(ctx->c, mesh->m, amg->a, w, bc);
            if (writeCache)
            {
                amg->a.signature = sig;
                amgSave(amg->a, sig, path);
            }
        }
        amg->a.signature = sig;
        if (how) *how = (int)r;
    });
}

int mfv_amg_clone(mfv_ctx ctx, mfv_amg src, mfv_amg* out)
{
    if (!src) return MFV_ERR_INVALID;
    return newAmg(ctx, src->mesh, out, [&](mfv_amg amg) {
        need(src->ctx == ctx, "amg clone: hierarchy of another context");
        amgBuildFrom(ctx->c, src->mesh->m, amg->a, src->a.agg, src->a.controls);
        amg->a.signature = src->a.signature;
    });
}

int mfv_amg_destroy(mfv_amg amg)
{
    if (!amg) return MFV_OK;
    const mfv_ctx ctx = amg->ctx;
    const mfv_mesh mesh = amg->mesh;
    (void)hipStreamSynchronize(ctx->c.stream);
    delete amg;
    releaseMesh(mesh);
    releaseCtx(ctx);
    return MFV_OK;
}

int mfv_amg_num_grids(mfv_amg amg, int* nGrids)
{
    if (!amg || !nGrids) return MFV_ERR_INVALID;
    *nGrids = (int)amg->a.grids.size();
    return MFV_OK;
}

int mfv_amg_grid_size(mfv_amg amg, int grid, int* nCells, int* nFaces)
{
    if (!amg || grid < 0 || grid >= (int)amg->a.grids.size()) return MFV_ERR_INVALID;
    if (nCells) *nCells = amg->a.grids[grid].n;
    if (nFaces) *nFaces = amg->a.grids[grid].nFaces;
    return MFV_OK;
}

int mfv_amg_coarse_mode(mfv_amg amg, int mode, int* tailStart, int* ldsGrid, int* ldsBytes)
{
    if (!amg || mode < -1 || mode > 2) return MFV_ERR_INVALID;
    if (mode >= 0) amg->a.coarseMode = mode;
    if (tailStart) *tailStart = amg->a.tailStart;
    if (ldsGrid) *ldsGrid = amg->a.tailLdsBytes ? amg->a.tailLdsGrid : 0;
    if (ldsBytes) *ldsBytes = amg->a.tailLdsBytes;
    return MFV_OK;
}

int mfv_amg_fine_stencil(mfv_amg amg, int* nOffsets, int* hOffsets)
{
    if (!amg || amg->a.grids.empty() || !nOffsets) return MFV_ERR_INVALID;
    const Grid& g = amg->a.grids[0];
    *nOffsets = g.diaM;
    if (hOffsets) std::copy(g.diaOff, g.diaOff + g.diaM, hOffsets);
    return MFV_OK;
}

int mfv_amg_level_addressing(mfv_amg amg, int level, int* hMap, int* hCoarseOwner, int* hCoarseNeighbour,
                             int* hFaceRestrict, signed char* hFaceFlip)
{
    if (!amg || level < 0 || level >= (int)amg->a.agg.size()) return MFV_ERR_INVALID;
    const AggLevel& L = amg->a.agg[level];
    if (hMap) std::copy(L.map.begin(), L.map.end(), hMap);
    if (hCoarseOwner) std::copy(L.cOwn.begin(), L.cOwn.end(), hCoarseOwner);
    if (hCoarseNeighbour) std::copy(L.cNei.begin(), L.cNei.end(), hCoarseNeighbour);
    if (hFaceRestrict) std::copy(L.faceRestrict.begin(), L.faceRestrict.end(), hFaceRestrict);
    if (hFaceFlip)
        for (size_t i = 0; i < L.faceFlip.size(); ++i) hFaceFlip[i] = (signed char)L.faceFlip[i];
    return MFV_OK;
}

namespace {

Prec precOf(int p)
{
    switch (p)
    {
    case MFV_PREC_FP64: return Prec::F64;
    case MFV_PREC_FP32: return Prec::F32;
    case MFV_PREC_FP16: return Prec::F16;
    }
    throw Error(kErrUnsupported, "mfv: precision " + std::to_string(p) + " is not one of 64, 32, 16");
}

} // namespace

int mfv_amg_galerkin(mfv_ctx ctx, mfv_amg amg, const double* dFineDiag, const double* dFineUpper,
                     const double* dFineLower, int precision)
{
    return guarded(ctx, [&] {
        need(ctx && amg && amg->ctx == ctx, "amg galerkin: bad context or hierarchy");
        need(dFineDiag && dFineUpper, "amg galerkin: fine matrix missing");
        if (dFineLower && dFineLower != dFineUpper)
            throw Error(kErrUnsupported, "mfv: amg galerkin: the hierarchy carries symmetric matrices only (one "
                                         "off-diagonal); pass lower == upper or NULL");
        amgGalerkin(ctx->c, amg->a, dFineDiag, dFineUpper, precOf(precision));
    });
}

int mfv_amg_grid_values(mfv_ctx ctx, mfv_amg amg, int grid, double* hDiag, double* hUpper)
{
    return guarded(ctx, [&] {
        need(ctx && amg && amg->ctx == ctx, "amg grid values: bad context or hierarchy");
        need(grid >= 1 && grid < (int)amg->a.grids.size(), "amg grid values: grid out of range");
        need(amg->a.galerkinDone, "amg grid values: no Galerkin yet");
        Grid& g = amg->a.grids[grid];
        hipCheck(hipStreamSynchronize(ctx->c.stream), "sync");
        if (hDiag) hipCheck(hipMemcpy(hDiag, g.diag64.p, sizeof(double) * g.n, hipMemcpyDeviceToHost), "diag");
        if (hUpper && g.nFaces)
            hipCheck(hipMemcpy(hUpper, g.upper64.p, sizeof(double) * g.nFaces, hipMemcpyDeviceToHost), "upper");
    });
}

int mfv_amg_vcycle(mfv_ctx ctx, mfv_amg amg, const double* dR, double* dZ, int nPreSweeps, int nPostSweeps,
                   int scaleCorrection)
{
    return guarded(ctx, [&] {
        need(ctx && amg && amg->ctx == ctx, "amg vcycle: bad context or hierarchy");
        need(dR && dZ, "amg vcycle: null vector");
        need(nPreSweeps >= 0 && nPostSweeps >= 0, "amg vcycle: negative sweeps");
        VcycleControls vc;
        vc.nPre = nPreSweeps;
        vc.nPost = nPostSweeps;
        vc.scale = scaleCorrection != 0;
        amgStartFrom(ctx->c, amg->a, dR, vc, nullptr);
        const void* res = amgVcycle(ctx->c, amg->a, vc, nullptr);
        amgResultToDouble(ctx->c, amg->a, res, dZ);
    });
}

int mfv_pcg_default_controls(mfv_pcg_controls* c)
{
    if (!c) return MFV_ERR_INVALID;
    c->tolerance = 1e-6;
    c->relTol = 0.0;
    c->maxIter = 1000;
    c->minIter = 0;
    c->precision = MFV_PREC_FP32;
    c->nPreSweeps = 1;
    c->nPostSweeps = 1;
    c->flexible = 0;
    c->smoother = 0;
    c->checkEvery = 4;
    c->nVcycles = 1;
    c->scaleCorrection = 0;
    c->captureVcycle = 0;
    return MFV_OK;
}

int mfv_dic_precondition(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* dR, double* dW)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(dR && dW && dR != dW, "dic: null or aliased argument");
        DBuf<double> rD(mesh->m.nCells);
        launchDicPrecondition(ctx->c, mesh->m, view(mesh, A), dR, dW, rD.p);
        hipCheck(hipStreamSynchronize(ctx->c.stream), "dic sync");
    });
}

int mfv_gauss_seidel(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* dB, double* dPsi, int nSweeps)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(dB && dPsi && dB != dPsi, "gauss-seidel: null or aliased argument");
        need(nSweeps >= 0, "gauss-seidel: negative sweep count");
        launchGaussSeidel(ctx->c, mesh->m, view(mesh, A), dB, dPsi, nSweeps);
    });
}

int mfv_sym_gauss_seidel(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* dB, double* dPsi, int nSweeps)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(dB && dPsi && dB != dPsi, "sym-gauss-seidel: null or aliased argument");
        need(nSweeps >= 0, "sym-gauss-seidel: negative sweep count");
        DBuf<double> bPrime(mesh->m.nCells);
        launchGaussSeidel(ctx->c, mesh->m, view(mesh, A), dB, dPsi, nSweeps, bPrime.p);
        hipCheck(hipStreamSynchronize(ctx->c.stream), "sym-gauss-seidel sync");
    });
}

int mfv_faceareapair_weights(int nFaces, const double* hSf, const double* hMagSf, double* hWeights)
{
    if (nFaces < 0 || (nFaces && (!hSf || !hMagSf || !hWeights))) return MFV_ERR_INVALID;
    for (int f = 0; f < nFaces; ++f)
    {
        const double root = std::sqrt(hMagSf[f]);
        const double wx = hSf[3 * f] / root;
        const double wy = (hSf[3 * f + 1] / root) * 1.01;
        const double wz = (hSf[3 * f + 2] / root) * 1.02;
        hWeights[f] = std::sqrt(wx * wx + wy * wy + wz * wz);
    }
    return MFV_OK;
}

int mfv_gamg_default_controls(mfv_gamg_controls* c)
{
    if (!c) return MFV_ERR_INVALID;
    const GamgControls d;
    c->smoother = (int)d.smoother;
    c->tolerance = d.tolerance;
    c->relTol = d.relTol;
    c->maxIter = d.maxIter;
    c->minIter = d.minIter;
    c->nPreSweeps = d.nPreSweeps;
    c->preSweepsLevelMultiplier = d.preSweepsLevelMultiplier;
    c->maxPreSweeps = d.maxPreSweeps;
    c->nPostSweeps = d.nPostSweeps;
    c->postSweepsLevelMultiplier = d.postSweepsLevelMultiplier;
    c->maxPostSweeps = d.maxPostSweeps;
    c->nFinestSweeps = d.nFinestSweeps;
    c->scaleCorrection = d.scaleCorrection ? 1 : 0;
    c->nVcycles = d.nVcycles;
    return MFV_OK;
}

int mfv_gamg_solve(mfv_ctx ctx, mfv_mesh mesh, mfv_amg amg, const mfv_ldu* A, const double* dB, double* dPsi,
                   const mfv_gamg_controls* controls, mfv_perf* perf)
{
    return guarded(ctx, [&] {
        const LduView v = gamgArgs(ctx, mesh, amg, A, dB, dPsi, controls);
        const PcgPerf r = gamgSolve(ctx->c, mesh->m, amg->a, v.diag, v.upper, dB, dPsi, gamgControls(*controls));
        putPerf(r, perf);
    });
}

int mfv_pcg_gamg_solve(mfv_ctx ctx, mfv_mesh mesh, mfv_amg amg, const mfv_ldu* A, const double* dB, double* dPsi,
                       double tolerance, double relTol, int maxIter, int minIter, const mfv_gamg_controls* controls,
                       mfv_perf* perf)
{
    return guarded(ctx, [&] {
        const LduView v = gamgArgs(ctx, mesh, amg, A, dB, dPsi, controls);
        const PcgPerf r = pcgGamgSolve(ctx->c, mesh->m, amg->a, v.diag, v.upper, dB, dPsi, tolerance, relTol, maxIter,
                                       minIter, gamgControls(*controls));
        putPerf(r, perf);
    });
}

int mfv_gamg_ref_solve(mfv_ctx ctx, mfv_mesh mesh, mfv_amg amg, const mfv_ldu* A, const double* dB, double* dPsi,
                       double tolerance, double relTol, int maxIter, int minIter, int nPostSweeps, int nFinestSweeps,
                       mfv_perf* perf)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(amg && amg->ctx == ctx && amg->mesh == mesh, "gamg.cu: hierarchy of another mesh or context");
        need(dB && dPsi && dB != dPsi, "gamg.cu: null or aliased argument");
        const LduView v = view(mesh, A);
        if (v.lower != v.upper) throw Error(kErrUnsupported, "mfv: gamg.cu: an asymmetric matrix is not carried");
        GamgRefControls g;
        g.tolerance = tolerance;
        g.relTol = relTol;
        g.maxIter = maxIter;
        g.minIter = minIter;
        g.nPostSweeps = nPostSweeps;
        g.nFinestSweeps = nFinestSweeps;
        putPerf(gamgRefSolve(ctx->c, mesh->m, amg->a, v.diag, v.upper, dB, dPsi, g), perf);
    });
}

int mfv_pcg_dic_solve(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* dB, double* dPsi,
                      const double* dNormFactor, const mfv_pcg_controls* controls, mfv_perf* perf)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(dB && dPsi && controls, "pcg-dic: null argument");
        PcgControls pc;
        pc.tolerance = controls->tolerance;
        pc.relTol = controls->relTol;
        pc.maxIter = controls->maxIter;
        pc.minIter = controls->minIter;
        pc.checkEvery = controls->checkEvery;
        const PcgPerf r = pcgDicSolve(ctx->c, mesh->m, view(mesh, A), dB, dPsi, dNormFactor, pc);
        if (perf)
        {
            perf->initialResidual = r.initialResidual;
            perf->finalResidual = r.finalResidual;
            perf->nIterations = r.nIterations;
            perf->converged = r.converged ? 1 : 0;
        }
    });
}

int mfv_amg_pcg_solve(mfv_ctx ctx, mfv_mesh mesh, mfv_amg amg, const mfv_ldu* A, const double* dB, double* dPsi,
                      const double* dNormFactor, const mfv_pcg_controls* controls, mfv_perf* perf)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(amg && amg->ctx == ctx && amg->mesh == mesh, "pcg: hierarchy of another mesh or context");
        need(dB && dPsi && controls, "pcg: null argument");
        const LduView v = view(mesh, A);
        if (v.lower != v.upper)
            throw Error(kErrUnsupported, "mfv: AMG-PCG solves symmetric matrices (lower == upper)");
        if (controls->smoother != 0)
            throw Error(kErrUnsupported, "mfv: pcg: smoother " + std::to_string(controls->smoother) +
                                             " is not carried (0 = weighted Jacobi)");
        if (controls->nVcycles > 1) throw Error(kErrUnsupported, "mfv: pcg: nVcycles above 1 is not carried");
        const PcgControls pc = pcgControls(*controls);
        const PcgPerf r = amgPcgSolve(ctx->c, mesh->m, amg->a, v.diag, v.upper, dB, dPsi, dNormFactor, pc);
        if (perf)
        {
            perf->initialResidual = r.initialResidual;
            perf->finalResidual = r.finalResidual;
            perf->nIterations = r.nIterations;
            perf->converged = r.converged ? 1 : 0;
        }
    });
}

int mfv_amg_pcg_solve_batch(mfv_ctx ctx, int n, const mfv_batch_solve* solves, mfv_perf* perfs)
{
    if (n < 0 || (n && !solves)) return MFV_ERR_INVALID;
    return guarded(ctx, [&] {
        std::vector<BatchSolve> b(n);
        for (int k = 0; k < n; ++k)
        {
            const mfv_batch_solve& q = solves[k];
            // the systems may live on other contexts of the same device (a serving slot's); the launch
            // goes on this context's stream, theirs must be idle meanwhile
            need(q.mesh && q.amg && q.amg->mesh == q.mesh, "pcg batch: null mesh, or a hierarchy of another mesh");
            need(q.amg->ctx->c.device == ctx->c.device, "pcg batch: a system on another device");
            need(q.dDiag && q.dB && q.dPsi, "pcg batch: null argument");
            if (q.controls.smoother != 0 || q.controls.nVcycles > 1)
                throw Error(kErrUnsupported, "mfv: pcg batch: weighted Jacobi and one cycle only");
            b[k].a = &q.amg->a;
            b[k].diag = q.dDiag;
            b[k].b = q.dB;
            b[k].psi = q.dPsi;
            b[k].pc = pcgControls(q.controls);
        }
        amgBatchSolve(ctx->c, b);
        if (perfs)
            for (int k = 0; k < n; ++k)
            {
                perfs[k].initialResidual = b[k].perf.initialResidual;
                perfs[k].finalResidual = b[k].perf.finalResidual;
                perfs[k].nIterations = b[k].perf.nIterations;
                perfs[k].converged = b[k].perf.converged ? 1 : 0;
            }
    });
}

int mfv_interpolate(mfv_ctx ctx, mfv_mesh mesh, const double* dWeights, const double* dVol, double* dFace,
                    double* dBFace)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(dWeights && dVol && dFace, "interpolate: null argument");
        launchInterpolate(ctx->c, mesh->m, dWeights, dVol, dFace, dBFace);
    });
}

int mfv_div_source(mfv_ctx ctx, mfv_mesh mesh, const double* dPhi, const double* dPhiB, const double* dV,
                   double sign, double* dSource)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(dPhi && dV && dSource, "div source: null argument");
        launchDivSource(ctx->c, mesh->m, dPhi, dPhiB, dV, sign, dSource);
    });
}

int mfv_flux_correct(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* dPsi, const double* dFfc,
                     const double* dPhiHbyA, double sign, double* dPhi)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(dPsi && dPhiHbyA && dPhi, "flux correct: null argument");
        launchFluxCorrect(ctx->c, mesh->m, view(mesh, A), dPsi, dFfc, dPhiHbyA, sign, dPhi);
    });
}

int mfv_fold_boundary_diag(mfv_ctx ctx, mfv_mesh mesh, const double* dDiag, const double* dIcX,
                           const double* dIcY, const double* dIcZ, double* dD)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(dDiag && dD && (mesh->m.nBFaces == 0 || (dIcX && dIcY && dIcZ)), "fold boundary diag: null argument");
        launchFoldBoundaryDiag(ctx->c, mesh->m, dIcX, dIcY, dIcZ, dDiag, dD);
    });
}

int mfv_reciprocal_v(mfv_ctx ctx, mfv_mesh mesh, const double* dD, const double* dV, double* dRAU)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(dD && dV && dRAU, "reciprocal V: null argument");
        launchReciprocalV(ctx->c, mesh->m, dD, dV, dRAU);
    });
}

int mfv_boundary_diag_term(mfv_ctx ctx, mfv_mesh mesh, const double* dIcX, const double* dIcY, const double* dIcZ,
                           int k, double* dBdDiag)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(k >= 0 && k < 3, "boundary diag term: component must be 0, 1 or 2");
        need(mesh->m.nBFaces == 0 || (dIcX && dIcY && dIcZ && dBdDiag), "boundary diag term: null argument");
        launchBoundaryDiagTerm(ctx->c, mesh->m.nBFaces, dIcX, dIcY, dIcZ, k, dBdDiag);
    });
}

int mfv_matrix_h(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* dPsi, const double* dSource,
                 const double* dBdDiag, const double* dBdSrc, const double* dV, const double* dRAU, int valid,
                 double* dOut)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(dOut != nullptr, "matrix H: null output");
        if (!valid)
        {
            launchZero(ctx->c, mesh->m.nCells, dOut);
            return;
        }
        const LduView L = view(mesh, A);
        need(dPsi && dSource && dV && (mesh->m.nBFaces == 0 || (dBdDiag && dBdSrc)), "matrix H: null argument");
        launchMatrixH(ctx->c, mesh->m, L.upper, L.lower, dPsi, dSource, dBdDiag, dBdSrc, dV, dRAU, dOut);
    });
}

int mfv_matrix_h3(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* const* dPsi,
                  const double* const* dSource, const double* const* dBdDiag, const double* const* dBdSrc,
                  const double* dV, const double* dRAU, const int* valid, double* const* dOut)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(dPsi && dSource && dBdDiag && dBdSrc && dOut && dV, "matrix H3: null argument");
        bool any = false;
        for (int k = 0; k < 3; ++k)
        {
            need(dOut[k] != nullptr, "matrix H3: null output");
            if (valid && !valid[k]) continue;
            any = true;
            need(dPsi[k] && dSource[k] && (mesh->m.nBFaces == 0 || (dBdDiag[k] && dBdSrc[k])),
                 "matrix H3: null argument");
        }
        if (!any)
        {
            for (int k = 0; k < 3; ++k) launchZero(ctx->c, mesh->m.nCells, dOut[k]);
            return;
        }
        const LduView L = view(mesh, A);
        launchMatrixH3(ctx->c, mesh->m, L.upper, L.lower, dPsi, dSource, dBdDiag, dBdSrc, dV, dRAU, valid, dOut);
    });
}

int mfv_relax_diag(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* dICbnd, double alpha,
                   const double* dCycSumOff, const double* dICmaxMag, const double* dICmin, double* dRelaxedDiag,
                   double* dDelta, int nComp, const double* const* dPsi, double* const* dSource)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(dRelaxedDiag || dDelta || nComp > 0, "relax: no output");
        need(nComp >= 0 && nComp <= 3, "relax: nComp must be 0..3");
        need(mesh->m.nBFaces == 0 || dICbnd || (dICmaxMag && dICmin), "relax: null internalCoeffs");
        if (nComp) need(dPsi && dSource, "relax: null psi/source");
        for (int k = 0; k < nComp; ++k) need(dPsi[k] && dSource[k], "relax: null psi/source");
        if (!(alpha > 0.0)) invalid("relax: alpha must be positive (fvMatrix::relax skips alpha <= 0)");
        const LduView L = view(mesh, A);
        launchRelaxDiag(ctx->c, mesh->m, L.upper, L.lower, L.diag, dICbnd, dICmaxMag, dICmin, dCycSumOff, alpha, nComp,
                        dPsi, dSource, dRelaxedDiag, dDelta);
    });
}

int mfv_linear_upwind_corr(mfv_ctx ctx, mfv_mesh mesh, const double* dPhi, const double* dCf, const double* dC,
                           int nComp, const double* const* dGrad, double* const* dCorr)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(nComp >= 1 && nComp <= 3, "linearUpwind correction: nComp must be 1..3");
        need(dGrad && dCorr, "linearUpwind correction: null gradient/correction");
        for (int k = 0; k < nComp; ++k) need(dGrad[k] && dCorr[k], "linearUpwind correction: null gradient/correction");
        need(mesh->m.nFaces == 0 || (dPhi && dCf && dC), "linearUpwind correction: null phi/Cf/C");
        if (mesh->m.nFaces == 0)
        {
            // no internal face: the correction is zero, and phi/Cf/C may be null
            for (int k = 0; k < nComp; ++k) launchZero(ctx->c, mesh->m.nCells, dCorr[k]);
            return;
        }
        launchLinearUpwindCorr(ctx->c, mesh->m, dPhi, dCf, dC, nComp, dGrad, dCorr);
    });
}

int mfv_scaled_dev2t(mfv_ctx ctx, int n, const double* dGradU, const double* dNu, double* dSigma)
{
    return guarded(ctx, [&] {
        need(n >= 0, "scaled dev2: negative size");
        need(n == 0 || (dGradU && dNu && dSigma), "scaled dev2: null argument");
        need(dGradU != dSigma, "scaled dev2: sigma must not alias gradU (the transpose reads across rows)");
        launchScaledDev2T(ctx->c, n, dGradU, dNu, dSigma);
    });
}

int mfv_tensor_div_source(mfv_ctx ctx, mfv_mesh mesh, const double* dWeights, const double* dSf,
                          const int* dIsEmpty, const double* dTcell, const double* dTbnd, double* dSrcX,
                          double* dSrcY, double* dSrcZ)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(dTcell && dSrcX && dSrcY && dSrcZ, "tensor div: null argument");
        need(mesh->m.nFaces + mesh->m.nBFaces == 0 || dSf, "tensor div: null Sf");
        need(mesh->m.nFaces == 0 || dWeights, "tensor div: null weights");
        need(mesh->m.nBFaces == 0 || dTbnd, "tensor div: null boundary tensors");
        launchTensorDiv(ctx->c, mesh->m, dWeights, dSf, dIsEmpty, dTcell, nullptr, dTbnd, dSrcX, dSrcY, dSrcZ);
    });
}

int mfv_div_dev_reff_explicit(mfv_ctx ctx, mfv_mesh mesh, const double* dWeights, const double* dSf,
                              const int* dIsEmpty, const double* dGradU, const double* dNu, const double* dSigmaB,
                              double* dSrcX, double* dSrcY, double* dSrcZ)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(dGradU && dNu && dSrcX && dSrcY && dSrcZ, "divDevReff: null argument");
        need(mesh->m.nFaces + mesh->m.nBFaces == 0 || dSf, "divDevReff: null Sf");
        need(mesh->m.nFaces == 0 || dWeights, "divDevReff: null weights");
        need(mesh->m.nBFaces == 0 || dSigmaB, "divDevReff: null boundary stress");
        launchTensorDiv(ctx->c, mesh->m, dWeights, dSf, dIsEmpty, dGradU, dNu, dSigmaB, dSrcX, dSrcY, dSrcZ);
    });
}

int mfv_cmpt_max_mag3(mfv_ctx ctx, int n, const double* dA, const double* dB, const double* dC, double* dOut)
{
    return guarded(ctx, [&] {
        need(n >= 0, "cmpt max mag: negative size");
        need(n == 0 || (dA && dB && dC && dOut), "cmpt max mag: null argument");
        launchCmpt3(ctx->c, n, dA, dB, dC, false, dOut);
    });
}

int mfv_cmpt_min3(mfv_ctx ctx, int n, const double* dA, const double* dB, const double* dC, double* dOut)
{
    return guarded(ctx, [&] {
        need(n >= 0, "cmpt min: negative size");
        need(n == 0 || (dA && dB && dC && dOut), "cmpt min: null argument");
        launchCmpt3(ctx->c, n, dA, dB, dC, true, dOut);
    });
}

int mfv_relax_field(mfv_ctx ctx, int n, double* dP, const double* dPPrev, double alpha)
{
    return guarded(ctx, [&] {
        need(n >= 0, "relax field: negative size");
        if (alpha >= 1.0 || alpha <= 0.0 || n == 0) return;
        need(dP && dPPrev, "relax field: null argument");
        launchRelaxField(ctx->c, n, dP, dPPrev, alpha);
    });
}

int mfv_constrain_hbya(mfv_ctx ctx, mfv_mesh mesh, const int* dTakeU, const double* dUb, const double* dHbyA,
                       double* dHbyAb)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(mesh->m.nBFaces == 0 || (dTakeU && dUb && dHbyA && dHbyAb), "constrain HbyA: null argument");
        launchConstrainHbyA(ctx->c, mesh->m, dTakeU, dUb, dHbyA, dHbyAb);
    });
}

int mfv_vector_flux(mfv_ctx ctx, mfv_mesh mesh, const double* dWeights, const double* dSf, const double* dUx,
                    const double* dUy, const double* dUz, double* dPhi)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(mesh->m.nFaces == 0 || (dWeights && dSf && dUx && dUy && dUz && dPhi), "vector flux: null argument");
        launchVectorFlux(ctx->c, mesh->m, dWeights, dSf, dUx, dUy, dUz, dPhi);
    });
}

int mfv_boundary_flux(mfv_ctx ctx, mfv_mesh mesh, const double* dSf, const double* dUxb, const double* dUyb,
                      const double* dUzb, const int* dIsEmpty, double* dPhiB)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(mesh->m.nBFaces == 0 || (dSf && dUxb && dUyb && dUzb && dPhiB), "boundary flux: null argument");
        launchBoundaryFlux(ctx->c, mesh->m, dSf ? dSf + 3 * (size_t)mesh->m.nFaces : nullptr, dUxb, dUyb, dUzb,
                           dIsEmpty, dPhiB);
    });
}

int mfv_adjust_phi(mfv_ctx ctx, int nBFaces, const int* dAdjustable, double* dPhiB, int nInternalFaces,
                   const double* dPhiInt, double* massCorr, int* closedVolume)
{
    return guarded(ctx, [&] {
        need(ctx != nullptr, "null context");
        need(nBFaces >= 0 && nInternalFaces >= 0, "adjust phi: negative size");
        need(nBFaces == 0 || (dAdjustable && dPhiB), "adjust phi: null argument");
        need(nInternalFaces == 0 || dPhiInt, "adjust phi: internal flux missing");
        Ctx& c = ctx->c;
        double* decision = c.scalars.p + 40;
        launchAdjustPhi(c, nBFaces, dAdjustable, dPhiB, nInternalFaces, dPhiInt, decision);
        hipCheck(hipMemcpyAsync(c.mailbox, decision, 3 * sizeof(double), hipMemcpyDeviceToHost, c.stream),
                 "adjust phi read-back");
        hipCheck(hipStreamSynchronize(c.stream), "adjust phi sync");
        const double mc = c.mailbox[0], refused = c.mailbox[1], closed = c.mailbox[2];
        if (massCorr) *massCorr = refused != 0.0 ? 1.0 : mc;
        if (closedVolume) *closedVolume = closed != 0.0 ? 1 : 0;
        if (refused != 0.0)
            invalid("adjust phi: the continuity error cannot be removed by adjusting the outflow (adjustPhi.C): "
                    "check the velocity boundary conditions");
    });
}

int mfv_grad_gauss(mfv_ctx ctx, mfv_mesh mesh, const int* dBcType, const double* dBcValue, const double* dSf,
                   const double* dWeights, const double* dV, const double* dPsi, double* dGrad)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(dSf && dV && dPsi && dGrad, "grad: null argument");
        need(!mesh->m.nFaces || dWeights, "grad: weights missing");
        need(!mesh->m.nBFaces || (dBcType && dBcValue), "grad: boundary types/values missing");
        launchGradGauss(ctx->c, mesh->m, dBcType, dBcValue, dSf, dWeights, dV, dPsi, dGrad);
    });
}

int mfv_laplacian_corr(mfv_ctx ctx, mfv_mesh mesh, const double* dGrad, const double* dWeights,
                       const double* dCorrVec, const double* dGammaf, const double* dMagSf, double* dFfc,
                       double* dSource)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(!mesh->m.nFaces || (dGrad && dWeights && dCorrVec && dGammaf && dMagSf && dFfc),
             "laplacian correction: null argument");
        launchLaplacianCorr(ctx->c, mesh->m, dGrad, dWeights, dCorrVec, dGammaf, dMagSf, dFfc, dSource);
    });
}

int mfv_laplacian_corr_limited(mfv_ctx ctx, mfv_mesh mesh, const double* dGrad, const double* dWeights,
                               const double* dCorrVec, const double* dGammaf, const double* dMagSf,
                               const double* dNonOrthDeltaCoeffs, const double* dVf, double limitCoeff,
                               double* dFfc, double* dSource)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(!mesh->m.nFaces || (dGrad && dWeights && dCorrVec && dGammaf && dMagSf && dFfc),
             "laplacian correction: null argument");
        const bool limited = limitCoeff > 0.0 && limitCoeff < 1.0;
        need(!limited || !mesh->m.nFaces || (dNonOrthDeltaCoeffs && dVf),
             "limited laplacian correction: nonOrthDeltaCoeffs and the field are needed");
        launchLaplacianCorr(ctx->c, mesh->m, dGrad, dWeights, dCorrVec, dGammaf, dMagSf, dFfc, dSource,
                            dNonOrthDeltaCoeffs, dVf, limitCoeff);
    });
}

int mfv_mesh_geometry(mfv_ctx ctx, mfv_mesh mesh, const double* dPoints, const int* dFaceOffsets,
                      const int* dFaceVerts, double* dCf, double* dSf, double* dMagSf, double* dC, double* dV,
                      double* dWeights, double* dDeltaCoeffs, double* dNonOrthDeltaCoeffs, double* dCorrVec,
                      double* dBDeltaCoeffs)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(dPoints && dFaceOffsets && dFaceVerts && dCf && dSf && dMagSf && dC && dV, "geometry: null argument");
        need(!mesh->m.nFaces || (dWeights && dDeltaCoeffs && dNonOrthDeltaCoeffs && dCorrVec),
             "geometry: interpolation outputs missing");
        need(!mesh->m.nBFaces || dBDeltaCoeffs, "geometry: boundary deltaCoeffs output missing");
        launchGeometry(ctx->c, mesh->m, dPoints, dFaceOffsets, dFaceVerts, dCf, dSf, dMagSf, dC, dV, dWeights,
                       dDeltaCoeffs, dNonOrthDeltaCoeffs, dCorrVec, dBDeltaCoeffs);
    });
}

int mfv_ctx_profile(mfv_ctx ctx, int enable)
{
    return guarded(ctx, [&] {
        need(ctx != nullptr, "null context");
        ctx->c.profiling = enable != 0;
        if ((int)ctx->c.prof.size() < kProfCount) ctx->c.prof.resize(kProfCount);
    });
}

int mfv_ctx_profile_read(mfv_ctx ctx, int tag, int* count, double* totalMs)
{
    return guarded(ctx, [&] {
        need(ctx != nullptr, "null context");
        need(tag >= 0 && tag < kProfCount, "profile: unknown tag");
        if ((int)ctx->c.prof.size() < kProfCount) ctx->c.prof.resize(kProfCount);
        hipCheck(hipStreamSynchronize(ctx->c.stream), "profile sync");
        double tot = 0.0;
        auto& v = ctx->c.prof[tag];
        for (auto& e : v)
        {
            float ms = 0.0f;
            hipCheck(hipEventElapsedTime(&ms, e.first, e.second), "event elapsed");
            tot += ms;
            ctx->c.eventPool.push_back(e.first);
            ctx->c.eventPool.push_back(e.second);
        }
        if (count) *count = (int)v.size();
        if (totalMs) *totalMs = tot;
        v.clear();
    });
}


int mfv_bicg_default_controls(mfv_bicg_controls* c)
{
    if (!c) return MFV_ERR_INVALID;
    c->tolerance = 1e-10;
    c->relTol = 0.0;
    c->maxIter = 1000;
    c->minIter = 0;
    c->preconditioner = MFV_PRECOND_DILU;
    c->checkEvery = 4;
    return MFV_OK;
}

} // extern "C"

namespace {

BicgControls bicgControls(const mfv_bicg_controls& c)
{
    if (c.preconditioner != MFV_PRECOND_DILU && c.preconditioner != MFV_PRECOND_DIAGONAL)
        throw Error(kErrUnsupported, "mfv: pbicgstab: preconditioner " + std::to_string(c.preconditioner) +
                                         " is not carried (DILU, diagonal)");
    if (c.maxIter < 0 || c.minIter < 0) invalid("pbicgstab: negative iteration limits");
    BicgControls b;
    b.tolerance = c.tolerance;
    b.relTol = c.relTol;
    b.maxIter = c.maxIter;
    b.minIter = c.minIter;
    b.precond = c.preconditioner == MFV_PRECOND_DILU ? BicgPrecond::DILU : BicgPrecond::Diagonal;
    b.checkEvery = c.checkEvery;
    return b;
}

PcgControls pcgControls(const mfv_pcg_controls& c)
{
    if (c.smoother != 0)
        throw Error(kErrUnsupported, "mfv: pcg: smoother " + std::to_string(c.smoother) +
                                         " is not carried (0 = weighted Jacobi)");
    if (c.nVcycles > 1) throw Error(kErrUnsupported, "mfv: pcg: nVcycles above 1 is not carried");
    PcgControls pc;
    pc.tolerance = c.tolerance;
    pc.relTol = c.relTol;
    pc.maxIter = c.maxIter;
    pc.minIter = c.minIter;
    pc.prec = precOf(c.precision);
    pc.nPre = c.nPreSweeps;
    pc.nPost = c.nPostSweeps;
    pc.flexible = c.flexible != 0;
    pc.scaleCorrection = c.scaleCorrection != 0;
    pc.checkEvery = c.checkEvery;
    pc.graph = c.captureVcycle != 0;
    return pc;
}

} // namespace

extern "C" {

int mfv_pbicgstab_solve(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* dSource, const double* dIC,
                        const double* dBC, double* dPsi, const mfv_bicg_controls* controls, mfv_perf* perf)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(dSource && dPsi && controls, "pbicgstab: null argument");
        need(mesh->m.nBFaces == 0 || (dIC && dBC), "pbicgstab: boundary coefficients missing");
        const LduView v = view(mesh, A);
        const PcgPerf r = pbicgstabSolve(ctx->c, mesh->m, v, dSource, dIC, dBC, dPsi, bicgControls(*controls));
        putPerf(r, perf);
    });
}

static_assert(MFV_STATIC_BICG_MAX_CELLS == kStaticBicgMaxCells, "the header states the static workspace's size");

int mfv_pbicgstab_solve_noalloc(mfv_ctx ctx, mfv_mesh mesh, const mfv_ldu* A, const double* dSource,
                                const double* dIC, const double* dBC, double* dPsi, const mfv_bicg_controls* controls,
                                mfv_perf* perf)
{
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(dSource && dPsi && controls, "pbicgstab (no allocation): null argument");
        need(mesh->m.nBFaces == 0 || (dIC && dBC), "pbicgstab (no allocation): boundary coefficients missing");
        const LduView v = view(mesh, A);
        const PcgPerf r = pbicgstabSolveStatic(ctx->c, mesh->m, v, dSource, dIC, dBC, dPsi, bicgControls(*controls));
        putPerf(r, perf);
    });
}

int mfv_simple_default_controls(mfv_simple_controls* c)
{
    if (!c) return MFV_ERR_INVALID;
    std::memset(c, 0, sizeof(*c));
    c->nu = 1e-5;
    c->relaxU = 0.7;
    c->relaxP = 0.3;
    c->momentumPredictor = 1;
    c->linearUpwind = 0;
    c->correctedLaplacian = 0;
    c->nNonOrthCorr = 0;
    c->pRefCell = -1;
    c->pRefValue = 0.0;
    mfv_bicg_default_controls(&c->U);
    mfv_pcg_default_controls(&c->p);
    c->p.tolerance = 1e-10;
    c->p.precision = MFV_PREC_FP64;
    c->mergeLevels = 1;
    c->pSolver = MFV_P_AMG_PCG;
    return MFV_OK;
}

int mfv_simple_create(mfv_ctx ctx, mfv_mesh mesh, const mfv_geometry* geometry, const int* hUType,
                      const double* hUValue, const int* hPType, const double* hPValue,
                      const mfv_simple_controls* controls, mfv_simple* out)
{
    if (out) *out = nullptr;
    return guarded(ctx, [&] {
        needMesh(ctx, mesh);
        need(out && geometry && controls, "simple create: null argument");
        const int nB = mesh->m.nBFaces;
        need(nB == 0 || (hUType && hUValue && hPType && hPValue), "simple create: boundary conditions missing");
        Geometry g;
        g.Sf = geometry->dSf;
        g.Cf = geometry->dCf;
        g.magSf = geometry->dMagSf;
        g.C = geometry->dC;
        g.V = geometry->dV;
        g.w = geometry->dWeights;
        g.dc = geometry->dDeltaCoeffs;
        g.ndc = geometry->dNonOrthDeltaCoeffs;
        g.corrVec = geometry->dCorrVec;
        g.bDelta = geometry->dBDeltaCoeffs;
        SimpleControls sc;
        sc.nu = controls->nu;
        sc.relaxU = controls->relaxU;
        sc.relaxP = controls->relaxP;
        sc.momentumPredictor = controls->momentumPredictor != 0;
        sc.linearUpwind = controls->linearUpwind != 0;
        sc.correctedLaplacian = controls->correctedLaplacian != 0;
        sc.nNonOrthCorr = controls->nNonOrthCorr;
        sc.pRefCell = controls->pRefCell;
        sc.pRefValue = controls->pRefValue;
        sc.U = bicgControls(controls->U);
        sc.P = pcgControls(controls->p);
        if (controls->pSolver != MFV_P_AMG_PCG && controls->pSolver != MFV_P_GAMG_REF)
            throw Error(kErrUnsupported, "mfv: simple: unknown pressure solver " + std::to_string(controls->pSolver));
        sc.pSolver = controls->pSolver == MFV_P_GAMG_REF ? PSolver::GamgRef : PSolver::AmgPcg;
        std::vector<int> ut(hUType, hUType + nB), pt(hPType, hPType + nB);
        std::vector<double> uv(hUValue, hUValue + 3 * (size_t)nB), pv(hPValue, hPValue + nB);
        std::unique_ptr<mfv_simple_s> h(new mfv_simple_s{ctx, mesh, nullptr});
        h->s.reset(new SimpleSolver(ctx->c, mesh->m, g, ut, uv, pt, pv, sc, controls->mergeLevels));
        retainCtx(ctx);
        retainMesh(mesh);
        *out = h.release();
    });
}

int mfv_simple_destroy(mfv_simple simple)
{
    if (!simple) return MFV_OK;
    const mfv_ctx ctx = simple->ctx;
    const mfv_mesh mesh = simple->mesh;
    (void)hipStreamSynchronize(ctx->c.stream);
    delete simple;
    releaseMesh(mesh);
    releaseCtx(ctx);
    return MFV_OK;
}

int mfv_simple_step(mfv_simple simple, const mfv_simple_fields* f, const double* dNuEff, const double* dNuEffB,
                    mfv_simple_result* result)
{
    if (!simple) return MFV_ERR_INVALID;
    return guarded(simple->ctx, [&] {
        need(f != nullptr, "simple step: null fields");
        SimpleFields sf;
        sf.U[0] = f->dUx;
        sf.U[1] = f->dUy;
        sf.U[2] = f->dUz;
        sf.Ub[0] = f->dUbx;
        sf.Ub[1] = f->dUby;
        sf.Ub[2] = f->dUbz;
        sf.p = f->dP;
        sf.pb = f->dPb;
        sf.phi = f->dPhi;
        sf.phiB = f->dPhiB;
        const SimpleResult r = simple->s->step(sf, dNuEff, dNuEffB);
        if (result)
        {
            for (int k = 0; k < 3; ++k)
            {
                result->UInitial[k] = r.UInitial[k];
                result->UFinal[k] = r.UFinal[k];
                result->UIters[k] = r.UIters[k];
            }
            result->pInitial = r.pInitial;
            result->pFinal = r.pFinal;
            for (int k = 0; k < 8; ++k) result->pIters[k] = r.pIters[k];
            result->massCorr = r.massCorr;
        }
    });
}

} // extern "C"

namespace {

mfv::Geometry geometryOf(const mfv_geometry* geometry)
{
    mfv::Geometry g;
    g.Sf = geometry->dSf;
    g.Cf = geometry->dCf;
    g.magSf = geometry->dMagSf;
    g.C = geometry->dC;
    g.V = geometry->dV;
    g.w = geometry->dWeights;
    g.dc = geometry->dDeltaCoeffs;
    g.ndc = geometry->dNonOrthDeltaCoeffs;
    g.corrVec = geometry->dCorrVec;
    g.bDelta = geometry->dBDeltaCoeffs;
    return g;
}

void putSimpleResult(const mfv::SimpleResult& r, mfv_simple_result* result)
{
    if (!result) return;
    for (int k = 0; k < 3; ++k)
    {
        result->UInitial[k] = r.UInitial[k];
        result->UFinal[k] = r.UFinal[k];
        result->UIters[k] = r.UIters[k];
    }
    result->pInitial = r.pInitial;
    result->pFinal = r.pFinal;
    for (int k = 0; k < 8; ++k) result->pIters[k] = r.pIters[k];
    result->massCorr = r.massCorr;
}

void putKeResult(const mfv::KEpsilonResult& r, mfv_kepsilon_result* out)
{
    if (!out) return;
    out->epsInitial = r.epsInitial;
    out->epsFinal = r.epsFinal;
    out->kInitial = r.kInitial;
    out->kFinal = r.kFinal;
    out->epsIters = r.epsIters;
    out->kIters = r.kIters;
    out->nWallCells = r.nWallCells;
}

void needKeFields(const mfv_kepsilon_fields* f, int nB)
{
    mfv::need(f != nullptr, "kepsilon: null fields");
    mfv::need(f->dK && f->dEps && f->dNut, "kepsilon: k, epsilon or nut missing");
    mfv::need(nB == 0 || (f->dKb && f->dEpsb && f->dNutb), "kepsilon: boundary values missing");
}

} // namespace

extern "C" {

int mfv_kepsilon_default_controls(mfv_kepsilon_controls* c)
{
    if (!c) return MFV_ERR_INVALID;
    std::memset(c, 0, sizeof(*c));
    c->nu = 1e-5;
    c->relaxEps = 1.0;
    c->relaxK = 1.0;
    mfv_bicg_default_controls(&c->solver);
    const mfv::KEpsilonCoeffs d;
    c->coeffs = mfv_kepsilon_coeffs{d.Cmu, d.C1, d.C2, d.C3, d.sigmaK, d.sigmaEps, d.kappa, d.E};
    return MFV_OK;
}

int mfv_near_wall_dist(int nInternalFaces, int nBFaces, const double* hPoints, const int* hFaceOffsets,
                       const int* hFaceVerts, const int* hBFaceCell, const double* hC, const int* hIsWall, double* hY)
{
    return mfv::guarded(nullptr, [&] {
        mfv::need(nInternalFaces >= 0 && nBFaces >= 0, "near wall dist: negative size");
        mfv::need(nBFaces == 0 || (hPoints && hFaceOffsets && hFaceVerts && hBFaceCell && hC && hIsWall && hY),
                  "near wall dist: null argument");
        mfv::nearWallDistHost(nInternalFaces, nBFaces, hPoints, hFaceOffsets, hFaceVerts, hBFaceCell, hC, hIsWall, hY);
    });
}

int mfv_kepsilon_create(mfv_ctx ctx, mfv_mesh mesh, const mfv_geometry* geometry, const int* hIsWall,
                        const double* hYWall, const int* hKType, const double* hKValue, const int* hEpsType,
                        const double* hEpsValue, const mfv_kepsilon_controls* controls, mfv_kepsilon* out)
{
    if (out) *out = nullptr;
    return mfv::guarded(ctx, [&] {
        needMesh(ctx, mesh);
        mfv::need(out && geometry && controls, "kepsilon create: null argument");
        const int nB = mesh->m.nBFaces;
        mfv::need(nB == 0 || (hIsWall && hYWall && hKType && hKValue && hEpsType && hEpsValue),
                  "kepsilon create: boundary arrays missing");
        mfv::KEpsilonControls kc;
        kc.nu = controls->nu;
        kc.relaxEps = controls->relaxEps;
        kc.relaxK = controls->relaxK;
        kc.solver = bicgControls(controls->solver);
        const mfv_kepsilon_coeffs& co = controls->coeffs;
        kc.co = mfv::KEpsilonCoeffs{co.Cmu, co.C1, co.C2, co.C3, co.sigmaK, co.sigmaEps, co.kappa, co.E};
        std::vector<int> w(hIsWall, hIsWall + nB), kt(hKType, hKType + nB), et(hEpsType, hEpsType + nB);
        std::vector<double> y(hYWall, hYWall + nB), kv(hKValue, hKValue + nB), ev(hEpsValue, hEpsValue + nB);
        std::unique_ptr<mfv_kepsilon_s> h(new mfv_kepsilon_s{ctx, mesh, nullptr, {}, {}});
        h->k.reset(new mfv::KEpsilon(ctx->c, mesh->m, geometryOf(geometry), w, y, kt, kv, et, ev, kc));
        retainCtx(ctx);
        retainMesh(mesh);
        *out = h.release();
    });
}

int mfv_kepsilon_destroy(mfv_kepsilon ke)
{
    if (!ke) return MFV_OK;
    const mfv_ctx ctx = ke->ctx;
    const mfv_mesh mesh = ke->mesh;
    (void)hipStreamSynchronize(ctx->c.stream);
    delete ke;
    releaseMesh(mesh);
    releaseCtx(ctx);
    return MFV_OK;
}

int mfv_kepsilon_nu_eff(mfv_kepsilon ke, const double* dNut, const double* dNutB, double* dNuEff, double* dNuEffB)
{
    if (!ke) return MFV_ERR_INVALID;
    return mfv::guarded(ke->ctx, [&] {
        const int nB = ke->mesh->m.nBFaces;
        mfv::need(dNut && dNuEff && (nB == 0 || (dNutB && dNuEffB)), "kepsilon nuEff: null argument");
        ke->k->nuEff(dNut, dNutB, dNuEff, dNuEffB);
    });
}

int mfv_kepsilon_correct(mfv_kepsilon ke, const mfv_kepsilon_fields* f, mfv_kepsilon_result* result)
{
    if (!ke) return MFV_ERR_INVALID;
    return mfv::guarded(ke->ctx, [&] {
        const int nB = ke->mesh->m.nBFaces;
        needKeFields(f, nB);
        mfv::need(f->dUx && f->dUy && f->dUz && f->dPhi, "kepsilon correct: U or phi missing");
        mfv::need(nB == 0 || (f->dUbx && f->dUby && f->dUbz && f->dPhiB), "kepsilon correct: boundary U or phi missing");
        const double* U[3] = {f->dUx, f->dUy, f->dUz};
        const double* Ub[3] = {f->dUbx, f->dUby, f->dUbz};
        putKeResult(ke->k->correct(U, Ub, f->dPhi, f->dPhiB, f->dK, f->dKb, f->dEps, f->dEpsb, f->dNut, f->dNutb),
                    result);
    });
}

int mfv_simple_step_kepsilon(mfv_simple simple, const mfv_simple_fields* f, mfv_kepsilon ke,
                             const mfv_kepsilon_fields* kf, mfv_simple_result* result, mfv_kepsilon_result* keResult)
{
    if (!simple || !ke) return MFV_ERR_INVALID;
    return mfv::guarded(simple->ctx, [&] {
        mfv::need(ke->ctx == simple->ctx && ke->mesh == simple->mesh,
                  "simple step kepsilon: the model and the step must share a context and a mesh");
        mfv::need(f != nullptr, "simple step: null fields");
        const int n = simple->mesh->m.nCells, nB = simple->mesh->m.nBFaces;
        needKeFields(kf, nB);
        ke->nuEffC.resize(std::max(n, 1));
        ke->nuEffB.resize(std::max(nB, 1));
        mfv::Ctx& c = simple->ctx->c;
        mfv::SimpleFields sf;
        sf.U[0] = f->dUx;
        sf.U[1] = f->dUy;
        sf.U[2] = f->dUz;
        sf.Ub[0] = f->dUbx;
        sf.Ub[1] = f->dUby;
        sf.Ub[2] = f->dUbz;
        sf.p = f->dP;
        sf.pb = f->dPb;
        sf.phi = f->dPhi;
        sf.phiB = f->dPhiB;
        {
            mfv::ProfScope ps(c, mfv::kProfSimpleTurbulence);
            ke->k->nuEff(kf->dNut, kf->dNutb, ke->nuEffC.p, ke->nuEffB.p);
        }
        putSimpleResult(simple->s->step(sf, ke->nuEffC.p, ke->nuEffB.p), result);
        const double* U[3] = {f->dUx, f->dUy, f->dUz};
        const double* Ub[3] = {f->dUbx, f->dUby, f->dUbz};
        mfv::KEpsilonResult kr;
        {
            mfv::ProfScope ps(c, mfv::kProfSimpleTurbulence);
            // the step's packed grad(U) is dead once its momentum equation is assembled: the model's
            // gradient of the corrected U goes there
            kr = ke->k->correct(U, Ub, f->dPhi, f->dPhiB, kf->dK, kf->dKb, kf->dEps, kf->dEpsb, kf->dNut, kf->dNutb,
                                simple->s->gradU.p);
        }
        putKeResult(kr, keResult);
    });
}

} // extern "C"

namespace mfv {

namespace {
hipEvent_t takeEvent(Ctx& c)
{
    if (!c.eventPool.empty())
    {
        hipEvent_t e = c.eventPool.back();
        c.eventPool.pop_back();
        return e;
    }
    // device-scope release: recording the event does not write the L2 back to memory, so the timed
    // window is the kernels' own
    hipEvent_t e;
    hipCheck(hipEventCreateWithFlags(&e, hipEventReleaseToDevice), "event create");
    return e;
}
} // namespace

ProfScope::ProfScope(Ctx& ctx, int t) : c(ctx), tag(t)
{
    if (!c.profiling || tag < 0 || tag >= kProfCount) return;
    e0 = takeEvent(c);
    e1 = takeEvent(c);
    hipCheck(hipEventRecord(e0, c.stream), "event record");
}

void ProfScope::close()
{
    if (!e0) return;
    hipEvent_t a = e0, b = e1;
    e0 = e1 = nullptr;
    if (hipEventRecord(b, c.stream) != hipSuccess) return;
    if ((int)c.prof.size() < kProfCount) c.prof.resize(kProfCount);
    c.prof[tag].emplace_back(a, b);
}

} // namespace mfv
