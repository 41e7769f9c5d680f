// Test infrastructure: a C-callable wrapper around the reference's own host C++ for the functions of the
// pressure path whose sources compile on their own (no CUDA headers): Amul, the OpenFOAM PCG with DIC,
// and one pairGAMGAgglomeration pass. Built by `make ref` into oracle/_ref/ from the reference's sources
// where they lie; used only by the tests, to check the C restatement (oracle.c) against the reference.
//
//   Amul                 src/OpenFOAM/matrices/ldu_spmv.cu:5-24
//   pcg                  src/OpenFOAM/matrices/pcg.cu:6-149
//   gamgPairAgglomerate  src/matrices/lduMatrix/solvers/GAMG/GAMGAgglomerations/pair_gamg_agglomeration_cpp.cu:158-284
//   gamgLduBand          same file, :347-365
//   pbicgstab            src/OpenFOAM/matrices/pbicgstab.cu:6-150 (PBiCGStab with DILU)
#include "ldu_spmv.cuh"
#include "pair_gamg_agglomeration_cpp.cuh"
#include "pbicgstab.cuh"
#include "pcg.cuh"

#include <exception>
#include <vector>

using namespace brae;

namespace {

// A mesh with only the addressing the solvers read: owner of every face (internal first, then the
// boundary faces), neighbour of the internal faces.
PrimitiveMesh ldu_mesh(int nCells, int nFaces, const int* owner, const int* neighbour, int nBFaces,
                       const int* bFaceCell)
{
    std::vector<label> own(owner, owner + nFaces), nei(neighbour, neighbour + nFaces);
    for (int i = 0; i < nBFaces; ++i) own.push_back(bFaceCell[i]);
    PrimitiveMesh m;
    m.assign({}, {}, std::vector<label>(own.size() + 1, 0), std::move(own), std::move(nei), {}, nCells);
    return m;
}

} // namespace

extern "C" {

int ref_amul(int nCells, int nFaces, const int* owner, const int* neighbour, const double* diag,
             const double* upper, const double* lower, const double* psi, double* out)
{
    try
    {
        const PrimitiveMesh m = ldu_mesh(nCells, nFaces, owner, neighbour, 0, nullptr);
        FvScalarMatrix M;
        M.diag.assign(diag, diag + nCells);
        M.upper.assign(upper, upper + nFaces);
        M.lower.assign(lower ? lower : upper, (lower ? lower : upper) + nFaces);
        const std::vector<scalar> x(psi, psi + nCells);
        const std::vector<scalar> y = Amul(M, x, m);
        for (int c = 0; c < nCells; ++c) out[c] = y[c];
        return 0;
    }
    catch (const std::exception&)
    {
        return 1;
    }
}

// PCG-DIC on a symmetric matrix whose boundary contributions arrive as one uncoupled patch
// (iC/bC per boundary face, bFaceCell its cells). psi in/out; perf = {initial, final, nIterations}.
int ref_pcg(int nCells, int nFaces, const int* owner, const int* neighbour, const double* diag,
            const double* upper, const double* source, int nBFaces, const int* bFaceCell, const double* iC,
            const double* bC, double* psi, double tolerance, double relTol, int maxIter, int minIter,
            double* perf)
{
    try
    {
        const PrimitiveMesh m = ldu_mesh(nCells, nFaces, owner, neighbour, nBFaces, bFaceCell);
        FvScalarMatrix M;
        M.diag.assign(diag, diag + nCells);
        M.upper.assign(upper, upper + nFaces);
        M.lower = M.upper;
        M.source.assign(source, source + nCells);
        std::vector<FvPatch> patches(1);
        patches[0].name = "boundary";
        patches[0].type = "patch";
        patches[0].start = nFaces;
        patches[0].size = nBFaces;
        patches[0].faceCells.assign(bFaceCell, bFaceCell + nBFaces);
        M.internalCoeffs.assign(1, std::vector<scalar>(iC, iC + nBFaces));
        M.boundaryCoeffs.assign(1, std::vector<scalar>(bC, bC + nBFaces));
        std::vector<scalar> x(psi, psi + nCells);
        const SolverPerformance sp = pcg(M, x, m, patches, tolerance, relTol, maxIter, minIter, nullptr);
        for (int c = 0; c < nCells; ++c) psi[c] = x[c];
        perf[0] = sp.initialResidual;
        perf[1] = sp.finalResidual;
        perf[2] = sp.nIterations;
        return 0;
    }
    catch (const std::exception&)
    {
        return 1;
    }
}

// PBiCGStab with DILU on an asymmetric matrix whose boundary arrives as one uncoupled patch (iC/bC per
// boundary face). psi in/out; perf = {initial, final, nIterations}.
int ref_pbicgstab(int nCells, int nFaces, const int* owner, const int* neighbour, const double* diag,
                  const double* upper, const double* lower, const double* source, int nBFaces, const int* bFaceCell,
                  const double* iC, const double* bC, double* psi, double tolerance, double relTol, int maxIter,
                  int minIter, double* perf)
{
    try
    {
        const PrimitiveMesh m = ldu_mesh(nCells, nFaces, owner, neighbour, nBFaces, bFaceCell);
        FvScalarMatrix M;
        M.diag.assign(diag, diag + nCells);
        M.upper.assign(upper, upper + nFaces);
        M.lower.assign(lower, lower + nFaces);
        M.source.assign(source, source + nCells);
        std::vector<FvPatch> patches(1);
        patches[0].name = "boundary";
        patches[0].type = "patch";
        patches[0].start = nFaces;
        patches[0].size = nBFaces;
        patches[0].faceCells.assign(bFaceCell, bFaceCell + nBFaces);
        M.internalCoeffs.assign(1, std::vector<scalar>(iC, iC + nBFaces));
        M.boundaryCoeffs.assign(1, std::vector<scalar>(bC, bC + nBFaces));
        std::vector<scalar> x(psi, psi + nCells);
        const SolverPerformance sp = pbicgstab(M, x, m, patches, tolerance, relTol, maxIter, minIter);
        for (int c = 0; c < nCells; ++c) psi[c] = x[c];
        perf[0] = sp.initialResidual;
        perf[1] = sp.finalResidual;
        perf[2] = sp.nIterations;
        return 0;
    }
    catch (const std::exception&)
    {
        return 1;
    }
}

// One pairGAMGAgglomeration pass over a level's addressing and face weights. forward in/out.
// restrictMap [nCells] out; returns the number of coarse cells, or -1 on an exception.
int ref_pair_agglomerate(int nCells, int nFaces, const int* lowerAddr, const int* upperAddr,
                         const double* faceWeights, int* forward, int* restrictMap)
{
    try
    {
        GamgLduAddressing a;
        a.nCells = nCells;
        a.lowerAddr.assign(lowerAddr, lowerAddr + nFaces);
        a.upperAddr.assign(upperAddr, upperAddr + nFaces);
        const std::vector<scalar> w(faceWeights, faceWeights + nFaces);
        bool fwd = *forward != 0;
        label nCoarse = 0;
        const std::vector<label> map = gamgPairAgglomerate(nCoarse, a, w, fwd);
        *forward = fwd ? 1 : 0;
        for (int c = 0; c < nCells; ++c) restrictMap[c] = map[c];
        return nCoarse;
    }
    catch (const std::exception&)
    {
        return -1;
    }
}

// lduAddressing::band() of a level: (bandwidth, profile)
int ref_ldu_band(int nCells, int nFaces, const int* lowerAddr, const int* upperAddr, int* bandwidth,
                 double* profile)
{
    try
    {
        GamgLduAddressing a;
        a.nCells = nCells;
        a.lowerAddr.assign(lowerAddr, lowerAddr + nFaces);
        a.upperAddr.assign(upperAddr, upperAddr + nFaces);
        const std::pair<label, scalar> b = gamgLduBand(a);
        *bandwidth = b.first;
        *profile = b.second;
        return 0;
    }
    catch (const std::exception&)
    {
        return 1;
    }
}

} // extern "C"
