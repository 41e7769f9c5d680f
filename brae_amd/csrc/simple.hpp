// One SIMPLE iteration on the device: the driver (simple.cpp) and the elementwise steps it composes
// (kernels_simple.hip). Reference: cpu::simpleStep, src/applications/solvers/simpleFoam/simpleFoam_cpp.cu:15-237.
#pragma once

#include "amg.hpp"

namespace mfv {

// boundary condition per boundary face (the C ABI's MFV_BC_*)
enum : int { kBcZeroGradient = 0, kBcFixedValue = 1, kBcEmpty = 2 };

void launchFill(Ctx& c, int n, double v, double* x);
void launchNegate(Ctx& c, int n, const double* a, double* out);
void launchAxpy(Ctx& c, int n, double s, const double* b, double* a);
void launchMomBoundary(Ctx& c, int nB, const int* type, const double* value, const double* phib, const double* nub,
                       const double* bMagSf, const double* bDelta, double* const* iC, double* const* bC);
void launchEvaluate(Ctx& c, const Mesh& m, int nComp, int stride, const int* type, const double* value,
                    const double* const* cell, double* const* out);
void launchPackGrad(Ctx& c, int n, const double* const* g, double* g9);
void launchStressBoundary(Ctx& c, const Mesh& m, const int* type, const double* Sf, const double* magSf,
                          const double* bDelta, const double* const* U, const double* const* Ub, const double* gradU,
                          const double* nub, double* sigB);
void launchStressSource(Ctx& c, int n, const double* V, const double* const* d, double* const* src);
void launchGradPSource(Ctx& c, int n, const double* gradP, const double* V, const double* const* src,
                       double* const* out);
void launchBoundaryFluxCorr(Ctx& c, const Mesh& m, const double* phiHbyAB, const double* iC, const double* bC,
                            const double* p, double* phiB);
void launchCorrectU(Ctx& c, int n, const double* rAU, const double* gradP, const double* const* HbyA, double* const* U);
void launchSubInPlace(Ctx& c, int n, const double* b, double* a);
void launchMul(Ctx& c, int n, const double* a, const double* b, double* y);

// geometry as launchGeometry writes it (Sf, Cf, magSf over internal then boundary faces)
struct Geometry
{
    const double *Sf = nullptr, *Cf = nullptr, *magSf = nullptr, *C = nullptr, *V = nullptr, *w = nullptr;
    const double *dc = nullptr, *ndc = nullptr, *corrVec = nullptr, *bDelta = nullptr;
};

enum class PSolver : int { AmgPcg = 0, GamgRef = 1 };

struct SimpleControls
{
    double nu = 1e-5;
    double relaxU = 0.7, relaxP = 0.3;
    bool momentumPredictor = true, linearUpwind = false, correctedLaplacian = false;
    int nNonOrthCorr = 0;
    int pRefCell = -1;
    double pRefValue = 0.0;
    BicgControls U;                 // the momentum components' PBiCGStab
    PcgControls P;                  // the pressure's AMG-PCG (GamgRef: its tolerance..minIter)
    PSolver pSolver = PSolver::AmgPcg;
};

struct SimpleResult
{
    double UInitial[3] = {}, UFinal[3] = {};
    int UIters[3] = {};
    double pInitial = 0, pFinal = 0;
    int pIters[8] = {};
    double massCorr = 1.0;
};

struct SimpleFields
{
    double *U[3] = {}, *Ub[3] = {}, *p = nullptr, *pb = nullptr, *phi = nullptr, *phiB = nullptr;
};

// The iteration's workspace and the pressure hierarchy, built once per mesh and case.
struct SimpleSolver
{
    Ctx& c;
    Mesh& m;
    Geometry g;
    SimpleControls ctl;
    DBuf<int> uType, pType, gType, pgType, isEmpty, takeU, adjustable;
    DBuf<double> uValue, pValue;
    Amg amg;
    // momentum matrix
    DBuf<double> D, Up, Lo, iC[3], bC[3], src[3], srcP[3];
    DBuf<double> lD, lU, nuC, nuB, nuCneg, nuF, nuFneg, nuBF;
    DBuf<double> g3[3], gradU, sigB, dv[3], corr[3], icMag, icMin;
    // pressure side
    DBuf<double> pPrev, gradP, Dm, rAU, HbyA[3], HbyAb[3], bd[3], phiHbyA, phiHbyAB;
    DBuf<double> rAUf, rAUb, pD, pU, piC, pbC, psrc, diagC, b, ffc;
    DBuf<double> decision;

    SimpleSolver(Ctx& ctx, Mesh& mesh, const Geometry& geom, const std::vector<int>& hUType,
                 const std::vector<double>& hUValue, const std::vector<int>& hPType, const std::vector<double>& hPValue,
                 const SimpleControls& controls, int mergeLevels);
    SimpleResult step(const SimpleFields& f, const double* nuEff, const double* nuEffB);
};

// ---- the standard k-epsilon model (kepsilon.hip; k_epsilon.cuh:221-337) ----
struct KEpsilonCoeffs
{
    double Cmu = 0.09, C1 = 1.44, C2 = 1.92, C3 = 0.0, sigmaK = 1.0, sigmaEps = 1.3, kappa = 0.41, E = 9.8;
};

struct KEpsilonControls
{
    double nu = 1e-5;
    double relaxEps = 1.0, relaxK = 1.0;
    BicgControls solver;   // the k and epsilon solves' PBiCGStab
    KEpsilonCoeffs co;
};

struct KEpsilonResult
{
    double epsInitial = 0, epsFinal = 0, kInitial = 0, kFinal = 0;
    int epsIters = 0, kIters = 0, nWallCells = 0;
};

struct KEpsilon
{
    Ctx& c;
    Mesh& m;
    Geometry g;
    KEpsilonControls ctl;
    int nWallCells = 0;
    DBuf<int> isWall, kType, epsType, isEmpty, gType, wallCell;
    DBuf<double> yWall, kValue, epsValue;
    DBuf<double> gradU, gbn, G, divU, D, Dr, Up, Lo, S, iC, bC, Df, lD, lU, nutB0;

    KEpsilon(Ctx& ctx, Mesh& mesh, const Geometry& geom, const std::vector<int>& hIsWall, const std::vector<double>& hY,
             const std::vector<int>& hKType, const std::vector<double>& hKValue, const std::vector<int>& hEpsType,
             const std::vector<double>& hEpsValue, const KEpsilonControls& controls);
    // nuEff = nu + nut in the cells and on the boundary faces
    void nuEff(const double* nut, const double* nutb, double* nuEffC, double* nuEffB);
    // gradScratch, when not null, is 9 nCells doubles whose contents the caller no longer needs (the SIMPLE
    // step's packed grad(U) once its momentum assembly is done); the gradient is built there instead of in
    // the model's own array, which is then never allocated
    KEpsilonResult correct(const double* const* U, const double* const* Ub, const double* phi, const double* phiB,
                           double* k, double* kb, double* eps, double* epsb, double* nut, double* nutb,
                           double* gradScratch = nullptr);

private:
    PcgPerf solveScalar(bool isEps, const double* phi, const double* phiB, const double* nut, double* psi,
                        double* psib, const double* prod, const double* eps, const double* k);
};

// nearWallDist (near_wall_dist.cuh:120-177) on the host: y per boundary face of the wall faces
void nearWallDistHost(int nFaces, int nBFaces, const double* points, const int* faceOffsets, const int* faceVerts,
                      const int* bFaceCell, const double* C, const int* isWall, double* y);

} // namespace mfv
