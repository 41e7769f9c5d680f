// Internal declarations shared by the mfv library's translation units: error handling, device
// buffers, the context (one HIP stream + scratch), the mesh's LDU addressing and the kernel
// launchers each .hip file provides. Nothing here crosses the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace mfv {

using label = int32_t;
using scalar = double;

struct Error : std::runtime_error
{
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum : int
{
    kOk = 0,
    kErrInvalid = 1,
    kErrHip = 2,
    kErrUnsupported = 3,
    kErrAlloc = 4,
    kErrState = 5,
};

inline void hipCheck(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        throw Error(kErrHip, std::string("mfv hip: ") + what + ": " + hipGetErrorString(e));
}

[[noreturn]] inline void invalid(const std::string& m) { throw Error(kErrInvalid, "mfv: " + m); }

// Device allocation owned by the library. Moves, never copies.
template <typename T>
struct DBuf
{
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { resize(count); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept
    {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void resize(size_t count)
    {
        if (count == n) return;
        release();
        if (count)
        {
            if (hipMalloc(&p, count * sizeof(T)) != hipSuccess)
            {
                p = nullptr;
                throw Error(kErrAlloc, "mfv: device allocation of " + std::to_string(count * sizeof(T)) + " bytes failed");
            }
        }
        n = count;
    }
    // Returns once the copy is complete: callers upload from pageable host arrays they free right after,
    // and a copy still in flight from freed host memory would read whatever the allocator puts there next
    // (another thread's data, in the serving loop). Uploads happen at set-up, not per iteration.
    void upload(const T* h, size_t count, hipStream_t s)
    {
        resize(count);
        if (count)
        {
            hipCheck(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s), "upload");
            hipCheck(hipStreamSynchronize(s), "upload sync");
        }
    }
    void upload(const std::vector<T>& h, hipStream_t s) { upload(h.data(), h.size(), s); }
    std::vector<T> download(hipStream_t s) const
    {
        std::vector<T> h(n);
        if (n)
        {
            hipCheck(hipMemcpyAsync(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost, s), "download");
            hipCheck(hipStreamSynchronize(s), "download sync");
        }
        return h;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Fixed partition of a reduction: a grid that depends on n only (two rows per thread), so every call
// on the same length sums in the same order. kMaxReduceBlocks partials per slot cover 134M rows; the
// finishing workgroup (kFinishBlock threads) reads them with several loads in flight per thread.
constexpr int kReduceBlock = 256;
constexpr int kMaxReduceBlocks = 1 << 18;
constexpr int kFinishBlock = 1024;

// Per-context scratch for two-stage reductions: partial sums per workgroup and the results (device
// scalars), plus a pinned host mailbox for read-backs.
struct Ctx
{
    int device = 0;
    hipStream_t stream = nullptr;
    bool ownsStream = false;
    std::string lastError;
    bool sequentialReduce = false;   // OpenFOAM order: one thread, ascending index
    DBuf<double> partials;           // kMaxReduceBlocks per reduction slot, 8 slots
    DBuf<double> scalars;            // 64 device scalars for solver state
    double* mailbox = nullptr;       // pinned host, 64 doubles
    int cuCount = 256;
    // kernel timing (off unless enabled): per tag, the event pairs recorded around its launches
    bool profiling = false;
    std::vector<std::vector<std::pair<hipEvent_t, hipEvent_t>>> prof;
    std::vector<hipEvent_t> eventPool;
};

// kernel classes the profiler can time
enum ProfTag : int
{
    kProfPcgSpmv64 = 0,   // p update + FP64 A.p + p.q (the outer product)
    kProfPcgUpdate,       // psi, r update + sum|r| + next cycle entry
    kProfPcgZr,           // z.r
    kProfVcycle,          // one whole V-cycle
    kProfAmgResidual0,    // grid-0 residual in the cycle precision
    kProfAmgCorrSmooth0,  // grid-0 correction update + first post-sweep (corrSmoothK or prolongSmoothK)
    kProfAmgRestrict0,    // grid 0 -> 1 restriction
    kProfAmgCorrAmul0,    // grid-0 scaled correction's A e and its two sums (corrAmulK)
    // the phases of one SIMPLE iteration (simple.cpp)
    kProfSimpleUEqn = 8,  // momentum matrix: convection, diffusion, linearUpwind and non-orthogonal
                          // corrections, the explicit stress, relax
    kProfSimpleUSolve,    // momentum predictor: -grad(p) source, the three component solves, U's patches
    kProfSimplePPredictor,// rAU, HbyA, constrainHbyA, phiHbyA, adjustPhi, interpolate(rAU)
    kProfSimplePAssembly, // per pressure pass: laplacian, boundary coefficients, correction, div, fold
    kProfSimpleGalerkin,  // the pressure hierarchy's coarse values
    kProfSimplePSolve,    // the pressure solves (AMG-PCG)
    kProfSimpleCorrect,   // phi, p relax, U = HbyA - rAU grad(p), patches
    kProfSimpleTurbulence,// the turbulence model's correct()
    kProfGridDown = 16,   // + grid: that grid's way down (sweeps, residual, restriction; the coarsest solve)
    kProfGridUp = 80,     // + grid: that grid's way up (prolongation, scaling, sweeps)
    kProfCount = 144
};

// Level schedule of the lower-triangular (forward) and upper-triangular (backward) dependency order of
// a mesh's cells: a cell's forward level is one past its highest lower neighbour's, so every cell of a
// level depends only on earlier levels. Cells are listed level by level, ascending within a level.
struct LevelSched
{
    std::vector<int> fwdStart, bwdStart;   // [levels + 1] into the lists (host)
    DBuf<int> fwdCells, bwdCells;          // [nCells] each
};

// PBiCGStab's work vectors, kept on the mesh between solves (pbicgstab.hip)
struct BicgWork;
struct BicgWorkDeleter
{
    void operator()(BicgWork* w) const;
};

// The mesh's internal-face LDU addressing on the device (lduAddressing), plus the boundary
// gather (cells' boundary faces in ascending face order).
struct Mesh
{
    int nCells = 0;
    int nFaces = 0;    // internal faces
    int nBFaces = 0;
    DBuf<label> owner, neighbour;
    DBuf<label> ownerStart;    // [nCells+1]
    DBuf<label> losortStart;   // [nCells+1]
    DBuf<label> losort;        // [nFaces] faces by neighbour, stable
    DBuf<label> losortOwner;   // [nFaces] owner[losort[k]]
    DBuf<label> bFaceCell;     // [nBFaces]
    DBuf<label> bndStart;      // [nCells+1] into bndFaces
    DBuf<label> bndFaces;      // boundary face indices (0-based within the boundary) per cell, ascending
    int band = 0;              // largest neighbour - owner over the internal faces
    std::vector<label> hOwner, hNeighbour, hBFaceCell;
    std::unique_ptr<LevelSched> sched;   // built on first use by the sequential sweeps (DIC, Gauss-Seidel)
    std::unique_ptr<BicgWork, BicgWorkDeleter> bicg;   // PBiCGStab's vectors, on first use
};

// the level schedule of the mesh's sequential sweeps, built once (dic.hip)
void buildLevelSchedule(Ctx& c, Mesh& m);

// ownerStart, stable losort by neighbour, boundary gather (abi.cpp)
void buildMesh(Ctx& c, Mesh& m, int nCells, int nFaces, const int* owner, const int* neighbour, int nB,
               const int* bFaceCell);

// ---- reductions / BLAS (kernels_blas.hip) ----
void launchDot(Ctx& c, int n, const double* a, const double* b, double* out, int slot);
void launchSumMag(Ctx& c, int n, const double* a, double* out, int slot);
void launchCopy(Ctx& c, int n, const double* src, double* dst);

// ---- LDU kernels (kernels_ldu.hip) ----
struct LduView
{
    int nCells;
    int nFaces;
    const double* diag;
    const double* upper;
    const double* lower;   // == upper for a symmetric matrix
};
void launchAmul(Ctx& c, const Mesh& m, const LduView& A, const double* x, double* y);
void launchResidual(Ctx& c, const Mesh& m, const LduView& A, const double* x, const double* b, double* r);
void launchNormFactor(Ctx& c, const Mesh& m, const LduView& A, const double* psi, const double* b, double* out);
void launchLaplacian(Ctx& c, const Mesh& m, const double* dc, const double* gammaf, const double* magSf,
                     double* diag, double* upper, double* lower);
void launchLaplacianBCoeffs(Ctx& c, int nB, const int* bcType, const double* bcValue, const double* gammab,
                            const double* bMagSf, const double* bDelta, double* iC, double* bC);
void launchDivUpwind(Ctx& c, const Mesh& m, const double* phi, double* diag, double* upper, double* lower);
void launchDivCentral(Ctx& c, const Mesh& m, const double* phi, const double* weights, double* diag, double* upper,
                      double* lower);
void launchDivBCoeffs(Ctx& c, int nB, const int* bcType, const double* bcValue, const double* phib, double* iC,
                      double* bC);
void launchFold(Ctx& c, const Mesh& m, const double* diag, const double* source, const double* iC, const double* bC,
                double* diagC, double* b);
void launchSetReference(Ctx& c, double* diag, double* b, int refCell, double refValue);
void launchFaceFlux(Ctx& c, const Mesh& m, const LduView& A, const double* psi, double* flux);
void launchInterpolate(Ctx& c, const Mesh& m, const double* w, const double* vol, double* face, double* bface);
void launchDivSource(Ctx& c, const Mesh& m, const double* phi, const double* phib, const double* V, double sign,
                     double* source);
void launchFluxCorrect(Ctx& c, const Mesh& m, const LduView& A, const double* psi, const double* ffc,
                       const double* phiHbyA, double sign, double* phi);
// ---- the pressure predictor's momentum glue (kernels_predictor.hip) ----
void launchZero(Ctx& c, int n, double* x);
void launchFoldBoundaryDiag(Ctx& c, const Mesh& m, const double* icX, const double* icY, const double* icZ,
                            const double* diag, double* D);
void launchReciprocalV(Ctx& c, const Mesh& m, const double* D, const double* V, double* rAU);
void launchBoundaryDiagTerm(Ctx& c, int nB, const double* icX, const double* icY, const double* icZ, int k,
                            double* out);
void launchMatrixH(Ctx& c, const Mesh& m, const double* upper, const double* lower, const double* psi,
                   const double* source, const double* bdDiag, const double* bdSrc, const double* V,
                   const double* rAU, double* out);
// the three components in one pass (pointers per component; valid may be null: all)
void launchMatrixH3(Ctx& c, const Mesh& m, const double* upper, const double* lower, const double* const* psi,
                    const double* const* source, const double* const* bdDiag, const double* const* bdSrc,
                    const double* V, const double* rAU, const int* valid, double* const* out);
void launchRelaxDiag(Ctx& c, const Mesh& m, const double* upper, const double* lower, const double* diag,
                     const double* icBnd, const double* icMag, const double* icMin, const double* cycSumOff,
                     double alpha, int nComp, const double* const* psi, double* const* source, double* relaxed,
                     double* delta);
void launchLinearUpwindCorr(Ctx& c, const Mesh& m, const double* phi, const double* Cf, const double* C, int nComp,
                            const double* const* grad, double* const* corr);
void launchScaledDev2T(Ctx& c, int n, const double* gradU, const double* nu, double* sigma);
void launchTensorDiv(Ctx& c, const Mesh& m, const double* w, const double* Sf, const int* isEmpty, const double* Tc,
                     const double* nu, const double* Tb, double* dX, double* dY, double* dZ);
void launchCmpt3(Ctx& c, int n, const double* x, const double* y, const double* z, bool min, double* out);
void launchRelaxField(Ctx& c, int n, double* p, const double* pPrev, double alpha);
void launchConstrainHbyA(Ctx& c, const Mesh& m, const int* takeU, const double* Ub, const double* HbyA,
                         double* HbyAb);
void launchVectorFlux(Ctx& c, const Mesh& m, const double* w, const double* Sf, const double* Ux, const double* Uy,
                      const double* Uz, double* phi);
void launchBoundaryFlux(Ctx& c, const Mesh& m, const double* SfB, const double* uxb, const double* uyb,
                        const double* uzb, const int* isEmpty, double* phiB);
// decision: 3 device doubles {massCorr, refused, closedVolume}
void launchAdjustPhi(Ctx& c, int nB, const int* adjustable, double* phiB, int nF, const double* phiInt,
                     double* decision);
// the gradients of three fields with the same patch types in one pass: each xyz per cell (g3[k], where not
// null) and, when g9 is given, packed component-major g9[(i*3 + k)*n + c] = d(psi_k)/dx_i
void launchGradGauss3(Ctx& c, const Mesh& m, const int* bcType, const double* const* bcValue, const double* Sf,
                      const double* w, const double* V, const double* const* psi, double* const* g3, double* g9);
void launchGradGauss(Ctx& c, const Mesh& m, const int* bcType, const double* bcValue, const double* Sf,
                     const double* w, const double* V, const double* psi, double* grad);
// limitCoeff in (0, 1): `limited k corrected` against nonOrthDc*(vf[nei] - vf[own]); otherwise unlimited
void launchLaplacianCorr(Ctx& c, const Mesh& m, const double* grad, const double* w, const double* corrVec,
                         const double* gammaf, const double* magSf, double* ffc, double* source,
                         const double* nonOrthDc = nullptr, const double* vf = nullptr, double limitCoeff = 0.0);

// ---- sequential-order sweeps on the level schedule (dic.hip) ----
struct PcgControls;
struct PcgPerf;
void launchDicPrecondition(Ctx& c, Mesh& m, const LduView& A, const double* r, double* w, double* rD);
// bPrime (n, may be null): symGaussSeidel, each sweep followed by the reverse walk from the forward walk's bPrime
void launchGaussSeidel(Ctx& c, Mesh& m, const LduView& A, const double* b, double* psi, int nSweeps,
                       double* bPrime = nullptr);
// DIC's reciprocal diagonal (calcReciprocalD)
void launchDicFactor(Ctx& c, Mesh& m, const LduView& A, double* rD);
// DICSmoother::smooth (gamg_solver_cpp.cu:277-293): per sweep psi += DIC(b - A psi); rA, w: n scratch each
void launchDicSmooth(Ctx& c, Mesh& m, const LduView& A, const double* rD, const double* b, double* psi, int nSweeps,
                     double* rA, double* w);
// guarded: OpenFOAM GAMG's coarsest-level PCG-DIC (gamg_solver_cpp.cu:321-383): stop without an update
// when |p.Ap|/normFactor < VSMALL, relTol counted only above 1e-20, no minIter
PcgPerf pcgDicSolve(Ctx& c, Mesh& m, const LduView& A, const double* b, double* psi, const double* normFactor,
                    const PcgControls& pc, bool guarded = false);

// ---- geometry (kernels_geom.hip) ----
void launchGeometry(Ctx& c, const Mesh& m, const double* points, const int* faceOffsets, const int* faceVerts,
                    double* Cf, double* Sf, double* magSf, double* C, double* V, double* weights, double* dc,
                    double* ndc, double* corr, double* bDelta);

// ---- kernel timing: HIP events around tagged launches on the context's stream ----
struct ProfScope
{
    Ctx& c;
    int tag;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ProfScope(Ctx& ctx, int t);
    ~ProfScope() { close(); }
    void close();   // record the end event now (idempotent)
};

} // namespace mfv
