// The AMG hierarchy on the device and the preconditioned-CG solver built on it.
//
// Layout. Every grid stores its off-diagonal coefficients ROW-WISE in a sliced-ELL layout: slices
// of 64 rows (one wavefront), each slice padded to its longest row, entries column-major inside the
// slice so lane l of a wavefront reads row 64*s + l and a wave's loads are contiguous. A row's entries
// are in ascending FACE order (its neighbour-side faces, then the faces it owns), so a row product
// sums in the order the reference's face loop does. Column indices are either explicit (int32 per
// entry) or, when the grid has at most 255 distinct row patterns (a structured hex mesh has 27), a
// one-byte pattern id per row that selects a table of column offsets -- index traffic drops from
// 4 bytes per entry to 1 byte per row.
//
// The FP64 face-form values of every coarse grid (upper per coarse face, diag per coarse cell) are
// kept as the Galerkin masters (brae's amgGalerkin, device_amg.cu:1693-1846); the sliced rows hold a
// copy in the V-cycle's precision (FP64, FP32, or FP16 scaled by a per-grid power of two).
#pragma once

#include "agglom.hpp"
#include "mfv_internal.hpp"

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace mfv {

constexpr int kSlice = 64;
constexpr int kDiaMax = 4;       // offsets of a fixed-stencil grid (a hex mesh has three)
constexpr double kOmega = 0.8;   // weighted-Jacobi relaxation (device_amg_detail.cuh:39)

enum class Prec : int { F64 = 64, F32 = 32, F16 = 16 };

// Sliced-ELL structure of one grid.
struct Sell
{
    int n = 0;
    int nSlices = 0;
    int width = 0;               // widest row
    long long nnz = 0;           // padded entries
    bool patterned = false;
    DBuf<int> slicePtr;          // [nSlices+1] first entry of each slice
    DBuf<int> cols;              // [nnz] explicit columns (padding: the row itself), when !patterned
    DBuf<uint8_t> pattern;       // [n] row pattern id, when patterned
    DBuf<int> patOff;            // [nPatterns*width] column offsets (padding 0)
    DBuf<int> slotLo;            // [nFaces] entry of face f in its owner's row (value upper[f])
    DBuf<int> slotHi;            // [nFaces] entry of face f in its neighbour's row (value lower[f])
};

struct Grid
{
    int n = 0;
    int nFaces = 0;
    Sell sell;
    // FP64 face-form values (coarse grids only; grid 0 reads the caller's matrix)
    DBuf<double> diag64, upper64;
    // V-cycle copies
    DBuf<double> off64;
    DBuf<float> off32;
    DBuf<__half> off16;
    DBuf<double> diagV64;
    DBuf<float> diagV32;
    // V-cycle vectors in the vector precision (double for the FP64 cycle, float otherwise)
    DBuf<double> b64, x64, y64, r64, e64;   // e: the coarse correction prolonged onto this grid
    DBuf<float> b32, x32, y32, r32, e32;
    DBuf<__half> b16, x16, y16, r16, e16;   // the FP16 cycle's (see the row engine's CycVec)
    // restriction/prolongation to the next grid (absent on the coarsest)
    DBuf<int> map;                       // [n] fine -> coarse
    DBuf<int> cellStart, cellList;       // coarse cell <- fine cells, ascending
    DBuf<int4> quads;                    // coarse cell <- its first four fine cells, ascending (-1: none), or
                                         // {-1, C, -1, -1} for an agglomerate of five or more (the lists)
    DBuf<int> dfaceStart, dfaceList;     // coarse cell <- fine faces inside it, ascending
    DBuf<int> faceStart, faceList;       // coarse face <- fine faces, ascending
    DBuf<int> denseOwn, denseNei;        // coarsest grid: addressing for the dense assembly
    // Grid 0 as diagonals, when its faces form a fixed stencil: every face's neighbour - owner is one of
    // diaM ascending positive offsets and the faces are in (owner, neighbour) order, so a row's entries in
    // face order are its columns ascending. The FP64 upper coefficient of face (r, r + diaOff[k]) sits at
    // dia64[k * diaPad + r] (0 where the face is absent); row r's lower entry at -diaOff[k] is the same
    // value at dia64[k * diaPad + r - diaOff[k]], so each coefficient is stored and streamed once.
    int diaM = 0;
    int diaOff[kDiaMax] = {};
    int diaPad = 0;                      // n rounded up to a slice
    DBuf<int> diaSlot;                   // [nFaces] k * diaPad + owner
    DBuf<double> dia64;                  // [diaM * diaPad]
    DBuf<float> dia32;                   // the V-cycle's copies in its coefficient precision (the FP16 one
    DBuf<__half> dia16;                  // divided by the grid's power-of-two scale, as off16)
    // host copies for inspection
    std::vector<int> hOwn, hNei;
};

struct AmgBuildControls
{
    AggKind kind = AggKind::Brae;
    int target = 64;
    int mergeLevels = 1;   // pairwise passes composed into each level
    bool fineRows = false; // grid 0 as sliced rows only, even on a fixed stencil (no diagonal copies)
};

// Work vectors and device state of the PCG solve on a hierarchy (FP64, one per cell), and the captured
// graph of a group of iterations with the key it was captured for.
struct PcgWork
{
    DBuf<double> r, p0, p1, q, rOld, st;   // r and rOld alternate as the residual and its predecessor
    DBuf<int> fl;
    DBuf<double> seqA, seqB;   // per-element terms for sequential (OpenFOAM-order) reductions
    hipGraphExec_t graph = nullptr;
    std::vector<uint64_t> graphKey;
    PcgWork() = default;
    PcgWork(const PcgWork&) = delete;
    PcgWork& operator=(const PcgWork&) = delete;
    ~PcgWork()
    {
        if (graph) (void)hipGraphExecDestroy(graph);
    }
};

// Device state of OpenFOAM's GAMG cycle on a hierarchy (of_gamg.hip): every coarse grid's LDU
// addressing, its correction and source, the DIC smoother's reciprocal diagonals, and the solve's
// FP64 work vectors.
struct OfWork
{
    std::vector<std::unique_ptr<Mesh>> mesh;        // [grid], grids >= 1
    std::vector<DBuf<double>> corr, src, rD;        // [grid]: coarseCorrFields, coarseSources, rD (grid 0: rD only)
    DBuf<double> ACf, preSmoothed, tR, tW, bPrime;  // fine-sized scratch
    DBuf<double> Apsi, fCorr, fRes, pA, wA, rA, AwA;
    DBuf<double> st;                                // device scalars
    DBuf<double> dense;                             // gamg.cu's coarsest elimination: n*n matrix + n rhs
};

struct Amg
{
    int nFine = 0;
    AmgBuildControls controls;
    uint64_t signature = 0;        // fingerprint of the inputs the structure was built from
    std::vector<AggLevel> agg;     // host addressing, agg[i]: grid i -> grid i+1
    std::vector<Grid> grids;       // grids[0] = the mesh
    int nCoarsest = 0;
    DBuf<double> coarseInv;        // dense inverse of the coarsest grid, row-major, FP64
    DBuf<double> coarseWork;       // Gauss-Jordan augmented matrix
    DBuf<double> scales;           // per grid: FP16 scale (1 otherwise)
    DBuf<double> alphas;           // per grid: the scaled correction's alpha
    DBuf<float> xmul;              // per grid: the FP16 cycle's solution-vector multiplier
    DBuf<double> sig;              // the FP16 cycle's input scale: [0] this cycle's, [1] the next one's
    DBuf<double> seqA, seqB;       // per-cell terms for sequential (OpenFOAM-order) reductions
    Prec prec = Prec::F32;         // precision the V-cycle copies are filled in
    int tailStart = 0;             // first grid the single-workgroup tail kernel runs (0: none)
    DBuf<uint8_t> tailDesc;        // per tail grid: pointers and sizes for the tail kernel
    int tailFor = -1;              // precision the descriptors were built for (-1: stale)
    int coarseMode = 2;            // 0: per-grid kernels; 1: the one-workgroup tail; 2: that tail with its
                                   // coarsest grids staged in LDS, as many as fit (else 1)
    DBuf<uint8_t> tailLdsBlob;     // the LDS-staged tail's slice starts, 16-bit columns and lists
    DBuf<uint8_t> tailLdsDesc;     // its descriptor
    int tailLdsBytes = 0;          // LDS it needs (0: the tail does not fit; tailK runs)
    int tailLdsGrid = 0;           // the first grid staged in LDS (the tail's grids before it are read in place)
    DBuf<uint8_t> cycleDesc;       // every grid's descriptor from grid 0 (batched whole solves)
    DBuf<double> batchWork;        // a batched solve's r, r_old, p, q
    bool galerkinDone = false;
    bool allocatedVectors = false;
    Prec vectorsFor = Prec::F32;
    std::unique_ptr<PcgWork> work;
    std::unique_ptr<OfWork> of;
};

// Build the hierarchy (host agglomeration + device structures).
void amgBuild(Ctx& c, const Mesh& m, Amg& a, const std::vector<double>& faceWeights, const AmgBuildControls& ctl);
// The device side of a build from given levels (a cache load or a clone); checks that the levels fit.
void amgBuildFrom(Ctx& c, const Mesh& m, Amg& a, std::vector<AggLevel> levels, const AmgBuildControls& ctl);

// Hierarchy persistence (amgHierarchySignature, device_amg.cu:1262-1285; writeAMGCache/readAMGCache,
// device_amg_cache.cu:176-326): a 64-bit fingerprint of the addressing, the face weights and the build
// controls; the structure (levels, no values) in one file written under a temporary name and renamed.
enum class CacheRead : int { Loaded = 0, Absent = 1, OtherMeshOrBuild = 2, Unreadable = 3 };
uint64_t amgSignature(const std::vector<int>& owner, const std::vector<int>& neighbour,
                      const std::vector<double>& weights, int nCells, const AmgBuildControls& ctl);
void amgSave(const Amg& a, uint64_t signature, const std::string& path);
// reads the levels when the file holds this signature; the caller then builds the device side
CacheRead amgReadLevels(const std::string& path, uint64_t signature, std::vector<AggLevel>& levels,
                        AmgBuildControls& ctl);
// Coarse matrices from the fine FP64 matrix (symmetric: lower == upper), V-cycle copies in `prec`,
// the coarsest inverse.
void amgGalerkin(Ctx& c, Amg& a, const double* fineDiag, const double* fineUpper, Prec prec);
// The FP64 coarse matrices alone (every grid's diag64/upper64), as OpenFOAM's GAMGSolver agglomerates
// its matrix (GAMGSolverAgglomerateMatrix.C): each coarse cell's diag is its fine cells' diags in
// ascending order, then twice each fine face inside it, ascending; each coarse face the sum of its fine
// faces, ascending.
void amgCoarseMatrices(Ctx& c, Amg& a, const double* fineDiag, const double* fineUpper);

struct VcycleControls
{
    int nPre = 1, nPost = 1;
    // Scale each grid's prolonged coarse correction e by the energy-minimising
    // alpha = (r.e)/(e.Ae) (brae's corrScaling, device_amg_vcycle.cu:1113-1126; OpenFOAM GAMG's scale),
    // fused with the first post-sweep: x = x + alpha*e + omega*(r - alpha*Ae)/D, r the residual the
    // grid restricted. Makes the preconditioner nonlinear: use the flexible CG.
    bool scale = false;
};

// One V-cycle on grid 0's b (set by amgStartFrom, with x = omega*b/D when nPre > 0). Returns the
// device buffer holding the result (double* for the FP64 cycle, float* otherwise). `stop` (device
// int, may be null) makes every kernel return at once when nonzero.
void* amgVcycle(Ctx& c, Amg& a, const VcycleControls& vc, const int* stop);

// Every device address a V-cycle's kernels take as an argument, hashed: a captured cycle stays valid
// while this is unchanged (values may change; buffers may not move).
uint64_t amgLayoutKey(const Amg& a);

// z (FP64) = the V-cycle result `res` (as amgVcycle returned it)
void amgResultToDouble(Ctx& c, Amg& a, const void* res, double* z);

// Fill grid 0's V-cycle vectors from an FP64 residual: b0 = r, x0 = omega*b0/D0 (pre-smooth from zero).
void amgStartFrom(Ctx& c, Amg& a, const double* r, const VcycleControls& vc, const int* stop);

// Row product on a grid's sliced rows in FP64 (the fine grid's FP64 copy): y = A x.
void sellAmul64(Ctx& c, const Grid& g, const double* diag, const double* x, double* y);

// ---- PCG ----
struct PcgControls
{
    double tolerance = 1e-6, relTol = 0.0;
    int maxIter = 1000, minIter = 0;
    Prec prec = Prec::F32;
    int nPre = 1, nPost = 1;
    bool flexible = false;
    bool scaleCorrection = false;
    int checkEvery = 4;
    bool graph = false;   // replay captured groups of iterations (captureVcycle)
};

struct PcgPerf
{
    double initialResidual = 0, finalResidual = 0;
    int nIterations = 0;
    bool converged = false;
};

// ---- PBiCGStab for asymmetric matrices (src/OpenFOAM/matrices/pbicgstab.cu:6-150; pbicgstab.hip) ----
enum class BicgPrecond : int { DILU = 0, Diagonal = 1 };
struct BicgControls
{
    double tolerance = 1e-10, relTol = 0.0;
    int maxIter = 1000, minIter = 0;
    BicgPrecond precond = BicgPrecond::DILU;
    int checkEvery = 4;
};
// the raw matrix (diag, upper, lower) with its uncoupled boundary coefficients per boundary face: the solve
// folds them (diag + iC, source + bC) as the reference's does; psi in place
PcgPerf pbicgstabSolve(Ctx& c, Mesh& m, const LduView& A, const double* source, const double* iC, const double* bC,
                       double* psi, const BicgControls& bc);
// The same solve with its workspace in a static device array instead of the mesh's buffers: no device
// allocation. Up to kStaticBicgMaxCells cells; DILU needs the mesh's level schedule built beforehand. Every
// call on one GPU shares that GPU's array, so calls that overlap on one GPU race.
constexpr int kStaticBicgMaxCells = 1 << 19;
void staticBicgScratchInit(int device);   // once per device, with that device current
PcgPerf pbicgstabSolveStatic(Ctx& c, Mesh& m, const LduView& A, const double* source, const double* iC,
                             const double* bC, double* psi, const BicgControls& bc);

// ---- OpenFOAM's GAMG (src/matrices/lduMatrix/solvers/GAMG/gamg_solver_cpp.cuh:67-85, GamgControls) ----
enum class OfSmoother : int { GaussSeidel = 0, SymGaussSeidel = 1, DIC = 2, DICGaussSeidel = 3 };
struct GamgControls
{
    OfSmoother smoother = OfSmoother::GaussSeidel;
    double tolerance = 1e-6, relTol = 0.0;
    int maxIter = 1000, minIter = 0;
    int nPreSweeps = 0, preSweepsLevelMultiplier = 1, maxPreSweeps = 4;
    int nPostSweeps = 2, postSweepsLevelMultiplier = 1, maxPostSweeps = 4;
    int nFinestSweeps = 2;
    bool scaleCorrection = true;
    int nVcycles = 2;   // per preconditioner application (GamgPreconditionerControls)
};

// GAMGSolver::solve on a folded symmetric system (gamgSolveFolded, gamg_solver_cpp.cu:979-1023):
// builds the FP64 coarse matrices from (diag, upper), then V-cycles until converged. psi in place.
// Synchronises once per V-cycle (the stopping test).
PcgPerf gamgSolve(Ctx& c, Mesh& m, Amg& a, const double* diag, const double* upper, const double* b, double* psi,
                  const GamgControls& g);
// The host CPU path's GAMG (src/OpenFOAM/matrices/gamg.cu:267-400, the pressure solver cpu::simpleStep
// calls): on a hierarchy built with AggKind::Pair on |Sf| down to nCellsInCoarsest, the folded symmetric
// system (diag, upper), psi in place. Per cycle the residual is restricted straight down, the coarsest level
// solved by Gaussian elimination with partial pivoting, then per level injection, gamg.cu's correction
// scaling and nPostSweeps Gauss-Seidel sweeps; the finest correction scaled against the residual, added, and
// psi smoothed nFinestSweeps times against b.
struct GamgRefControls
{
    double tolerance = 1e-6, relTol = 0.0;
    int maxIter = 1000, minIter = 0;
    int nPostSweeps = 2, nFinestSweeps = 2;
};
PcgPerf gamgRefSolve(Ctx& c, Mesh& m, Amg& a, const double* diag, const double* upper, const double* b, double* psi,
                     const GamgRefControls& g);
// PCG with the GAMG preconditioner (pcgGamgSolve, gamg_solver_cpp.cu:1027-1143): each application runs
// g.nVcycles cycles from zero; tolerance..minIter are the PCG's, g's tolerance/relTol the coarsest
// solve's. Synchronises once per iteration.
PcgPerf pcgGamgSolve(Ctx& c, Mesh& m, Amg& a, const double* diag, const double* upper, const double* b, double* psi,
                     double tolerance, double relTol, int maxIter, int minIter, const GamgControls& g);

// Whole AMG-PCG solves batched into one launch, one workgroup each (each result is the same bits in any
// batch). Every hierarchy's Galerkin values in its solve's precision.
struct BatchSolve
{
    Amg* a = nullptr;
    const double* diag = nullptr;   // the folded FP64 diagonal (the upper is the one Galerkin saw)
    const double* b = nullptr;
    double* psi = nullptr;          // initial guess in, solution out
    PcgControls pc;
    PcgPerf perf;                   // out
};
void amgBatchSolve(Ctx& c, std::vector<BatchSolve>& solves);

// AMG-preconditioned CG on the folded fine matrix (diag, upper symmetric). psi updated in place.
// normFactor: device scalar, or null to compute OpenFOAM's normFactor here.
PcgPerf amgPcgSolve(Ctx& c, const Mesh& m, Amg& a, const double* diag, const double* upper, const double* b,
                    double* psi, const double* normFactor, const PcgControls& pc);

} // namespace mfv
