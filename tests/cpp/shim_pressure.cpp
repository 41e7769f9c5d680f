// A C++ caller of the pressure path through brae_mfv.hpp, written the way brae's own drivers call it
// (simpleFoam.cu:343-381): assemble -laplacian on a box, fold, reference cell, buildAMG, amgGalerkin,
// deviceAMGPCG, residual check with deviceAmul and deviceNormFactorInto, flux; the hierarchy persisted,
// reloaded and cloned (buildOrLoadAMG, cloneAMG); OpenFOAM's GAMG on the same system
// (faceAreaPairGamgAgglomeration, gamgSolveFolded, pcgGamgSolveFolded, and an fvSolution entry read by
// readGamgControls); and the error
// behaviour (refusals throw std::runtime_error("brae hip: ...")). Prints one JSON line; exit status 0
// when every check holds.
#include "brae_foam.hpp"
#include "brae_mfv.hpp"

#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

using namespace brae;

namespace {

struct Box
{
    int nC = 0;
    std::vector<label> owner, neighbour, faceOffsets, faceVerts;
    std::vector<scalar> points;
};

// nx*ny*nz hex box of the given lengths: internal faces by owner (+x, +y, +z neighbours), then the six
// boundary patches (x-min, x-max, y-min, y-max, z-min, z-max), every face a quad with its normal out of
// the owner; interior points jittered by +-jit of the local spacing
Box makeBox(int nx, int ny, int nz, double lx, double ly, double lz, double jit, unsigned seed)
{
    Box b;
    b.nC = nx * ny * nz;
    const int px = nx + 1, py = ny + 1, pz = nz + 1;
    auto P = [&](int i, int j, int k) { return i + px * (j + py * k); };
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    b.points.resize(3 * (size_t)px * py * pz);
    for (int k = 0; k < pz; ++k)
        for (int j = 0; j < py; ++j)
            for (int i = 0; i < px; ++i)
            {
                double x = lx * i / nx, y = ly * j / ny, z = lz * k / nz;
                if (i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz)
                {
                    x += jit * lx / nx * u(rng);
                    y += jit * ly / ny * u(rng);
                    z += jit * lz / nz * u(rng);
                }
                const int p = P(i, j, k);
                b.points[3 * p] = x;
                b.points[3 * p + 1] = y;
                b.points[3 * p + 2] = z;
            }
    auto quad = [&](int a, int bb, int c, int d) {
        b.faceVerts.insert(b.faceVerts.end(), {a, bb, c, d});
    };
    for (int c = 0; c < b.nC; ++c)
    {
        const int i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
        if (i < nx - 1)
        {
            b.owner.push_back(c);
            b.neighbour.push_back(c + 1);
            quad(P(i + 1, j, k), P(i + 1, j + 1, k), P(i + 1, j + 1, k + 1), P(i + 1, j, k + 1));
        }
        if (j < ny - 1)
        {
            b.owner.push_back(c);
            b.neighbour.push_back(c + nx);
            quad(P(i, j + 1, k), P(i, j + 1, k + 1), P(i + 1, j + 1, k + 1), P(i + 1, j + 1, k));
        }
        if (k < nz - 1)
        {
            b.owner.push_back(c);
            b.neighbour.push_back(c + nx * ny);
            quad(P(i, j, k + 1), P(i + 1, j, k + 1), P(i + 1, j + 1, k + 1), P(i, j + 1, k + 1));
        }
    }
    for (int side = 0; side < 6; ++side)
        for (int c = 0; c < b.nC; ++c)
        {
            const int i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
            switch (side)
            {
            case 0: if (i == 0) { b.owner.push_back(c); quad(P(i, j, k), P(i, j, k + 1), P(i, j + 1, k + 1), P(i, j + 1, k)); } break;
            case 1: if (i == nx - 1) { b.owner.push_back(c); quad(P(i + 1, j, k), P(i + 1, j + 1, k), P(i + 1, j + 1, k + 1), P(i + 1, j, k + 1)); } break;
            case 2: if (j == 0) { b.owner.push_back(c); quad(P(i, j, k), P(i + 1, j, k), P(i + 1, j, k + 1), P(i, j, k + 1)); } break;
            case 3: if (j == ny - 1) { b.owner.push_back(c); quad(P(i, j + 1, k), P(i, j + 1, k + 1), P(i + 1, j + 1, k + 1), P(i + 1, j + 1, k)); } break;
            case 4: if (k == 0) { b.owner.push_back(c); quad(P(i, j, k), P(i, j + 1, k), P(i + 1, j + 1, k), P(i + 1, j, k)); } break;
            case 5: if (k == nz - 1) { b.owner.push_back(c); quad(P(i, j, k + 1), P(i + 1, j, k + 1), P(i + 1, j + 1, k + 1), P(i, j + 1, k + 1)); } break;
            }
        }
    for (size_t f = 0; f <= b.owner.size(); ++f) b.faceOffsets.push_back(static_cast<label>(4 * f));
    return b;
}

template <typename F>
bool throwsBrae(F f)
{
    try
    {
        f();
    }
    catch (const std::runtime_error& e)
    {
        return std::string(e.what()).rfind("brae hip: ", 0) == 0;
    }
    return false;
}

} // namespace

int main(int argc, char** argv)
{
    const int n = argc > 1 ? std::atoi(argv[1]) : 32;
    const Box box = makeBox(n, n, n, 1.0, 1.0, 1.0, 0.1, 7);
    DeviceMesh dm;
    dm.build(box.nC, box.owner, box.neighbour, box.points, box.faceOffsets, box.faceVerts);
    const int nC = dm.nCells, nI = dm.nInternalFaces, nB = dm.nBoundaryFaces;

    // the pressure matrix the way brae holds it (negated laplacian, positive definite): gamma_f = -1
    DeviceBuffer<scalar> gammaf(std::vector<scalar>(nI, -1.0));
    DeviceBuffer<scalar> diag, upper;
    deviceLaplacianCoeffsSymmetric(dm, gammaf, diag, upper, true);
    // all-Neumann pressure with a zero-mean source and a reference cell
    std::mt19937_64 rng(3);
    std::normal_distribution<double> g(0.0, 1.0);
    std::vector<scalar> src(nC);
    double mean = 0.0;
    for (auto& s : src) mean += (s = g(rng));
    mean /= nC;
    for (auto& s : src) s -= mean;
    DeviceBuffer<scalar> source(src), iC(std::vector<scalar>(nB, 0.0)), bC(std::vector<scalar>(nB, 0.0));
    DeviceBuffer<scalar> diagC, b;
    deviceFold(dm, diag, source, iC, bC, diagC, b);
    deviceSetReference(diagC, b, 0, 0.0);

    std::vector<label> own(box.owner.begin(), box.owner.begin() + nI);
    const std::vector<scalar> magSf = dm.magSfInternal.download();
    AMGData amg = buildAMG(own, box.neighbour, magSf, nC);
    amgGalerkin(amg, diagC, upper, upper);
    const DeviceLduView A = lduView(dm, diagC, upper);
    DeviceBuffer<scalar> psi(std::vector<scalar>(nC, 0.0));
    const DeviceSolverPerf perf = deviceAMGPCG(A, amg, b, psi, static_cast<const scalar*>(nullptr), 1e-8, 0.0, 500,
                                               false, 4, true, 0);

    // independent residual check: sum|b - A psi| / normFactor(psi, b)
    DeviceBuffer<scalar> Apsi, nf, ones;
    deviceAmul(A, psi, Apsi);
    deviceNormFactorInto(A, psi, b, ones, nf);
    const std::vector<scalar> hA = Apsi.download(), hb = b.download(), hnf = nf.download();
    double res = 0.0;
    for (int c = 0; c < nC; ++c) res += std::fabs(hb[c] - hA[c]);
    res /= hnf[0];
    DeviceBuffer<scalar> flux;
    deviceMatrixFluxInternal(A, psi, flux);

    // refusals by name
    const bool refusesSA = throwsBrae([&] {
        const bool sa = true;
        AMGData x = buildAMG(own, box.neighbour, magSf, nC, &sa);
    });
    DeviceBuffer<scalar> lower(std::vector<scalar>(nI, -0.5));
    const bool refusesAsym = throwsBrae([&] { amgGalerkin(amg, diagC, upper, lower); });

    // OpenFOAM's GAMG on the same folded system: as the solver and as the PCG's preconditioner
    const std::vector<scalar> SfAll = dm.Sf.download();
    const std::vector<scalar> SfI(SfAll.begin(), SfAll.begin() + 3 * (size_t)nI);
    AMGData gamg = faceAreaPairGamgAgglomeration(own, box.neighbour, SfI, magSf, nC, 10);
    GamgControls gc;
    gc.tolerance = 1e-8;
    DeviceBuffer<scalar> psiG(std::vector<scalar>(nC, 0.0)), psiP(std::vector<scalar>(nC, 0.0));
    const SolverPerformance pg = gamgSolveFolded(gamg, diagC, upper, b, psiG, gc);
    const SolverPerformance pp = pcgGamgSolveFolded(gamg, diagC, upper, b, psiP, 1e-8, 0.0, 500, 0,
                                                    GamgPreconditionerControls{});
    const bool refusesSmoother = throwsBrae([&] {
        GamgControls x;
        x.smoother = "nonBlockingGaussSeidel";
        DeviceBuffer<scalar> t(std::vector<scalar>(nC, 0.0));
        gamgSolveFolded(gamg, diagC, upper, b, t, x);
    });

    // the pressure predictor's glue on an upwind momentum matrix: H()'s defining identity
    // V*H = source + bdDiag*psi + bdSrc - (off-diagonal A) psi against deviceAmul, the flux of a uniform
    // field (U . Sf), and adjustPhi closing continuity
    std::vector<scalar> hPhi(nI);
    for (auto& v : hPhi) v = g(rng);
    DeviceBuffer<scalar> phiM(hPhi), mDiag, mUpper, mLower;
    deviceDivUpwindCoeffs(dm, phiM, mDiag, mUpper, mLower);
    const DeviceLduView M = lduView(dm, mDiag, mUpper, &mLower);
    DeviceBuffer<scalar> iCs[3] = {DeviceBuffer<scalar>(std::vector<scalar>(nB, 0.3)),
                                   DeviceBuffer<scalar>(std::vector<scalar>(nB, 0.5)),
                                   DeviceBuffer<scalar>(std::vector<scalar>(nB, 0.7))};
    DeviceBuffer<scalar> Dm, rAU, bd, Hk, HbyAk, zeroDiag(std::vector<scalar>(nC, 0.0)), offPsi;
    deviceFoldBoundaryDiag(dm, mDiag, iCs, Dm);
    deviceReciprocalV(dm, Dm, rAU);
    deviceBoundaryDiagTerm(dm, iCs, 1, bd);
    std::vector<scalar> hU(nC), hBc(nB);
    for (auto& v : hU) v = g(rng);
    for (auto& v : hBc) v = g(rng);
    DeviceBuffer<scalar> Uk(hU), bcK(hBc);
    deviceMatrixH(M, dm, Uk, source, bd, bcK, Hk);
    deviceMatrixHbyA(M, dm, Uk, source, bd, bcK, rAU, HbyAk);
    deviceAmul(lduView(dm, zeroDiag, mUpper, &mLower), Uk, offPsi);
    const std::vector<scalar> hH = Hk.download(), hHb = HbyAk.download(), hOff = offPsi.download(),
                              hV = dm.V.download(), hR = rAU.download(), hBd = bd.download();
    std::vector<scalar> rhs(src);
    for (int i = 0; i < nB; ++i) rhs[box.owner[nI + i]] += hBd[i] * hU[box.owner[nI + i]] + hBc[i];
    double hErr = 0.0, hScale = 0.0, hbErr = 0.0;
    for (int c = 0; c < nC; ++c)
    {
        hErr = std::max(hErr, std::fabs(hV[c] * hH[c] - (rhs[c] - hOff[c])));
        hScale = std::max(hScale, std::fabs(rhs[c]) + std::fabs(hOff[c]));
        hbErr = std::max(hbErr, std::fabs(hHb[c] - hR[c] * hH[c]));
    }
    DeviceBuffer<scalar> u1(std::vector<scalar>(nC, 0.25)), u2(std::vector<scalar>(nC, -0.5)),
        u3(std::vector<scalar>(nC, 1.0)), phiU;
    deviceVectorFlux(dm, u1, u2, u3, phiU);
    const std::vector<scalar> hPhiU = phiU.download();
    double fluxErr = 0.0;
    for (int f = 0; f < nI; ++f)
        fluxErr = std::max(fluxErr, std::fabs(hPhiU[f] - (0.25 * SfAll[3 * f] - 0.5 * SfAll[3 * f + 1] +
                                                          SfAll[3 * f + 2])));
    std::vector<scalar> hPb(nB);
    std::vector<label> hAdj(nB);
    for (int i = 0; i < nB; ++i)
    {
        hPb[i] = g(rng);
        hAdj[i] = (i % 3 == 0) ? 1 : 0;
    }
    DeviceBuffer<scalar> phiB(hPb);
    DeviceBuffer<label> adjustable(hAdj);
    bool closedVolume = true;
    deviceAdjustPhi(adjustable, phiB, &phiU, &closedVolume);
    const std::vector<scalar> hPa = phiB.download();
    double net = 0.0, tot = 0.0;
    for (int i = 0; i < nB; ++i)
    {
        net += hPa[i];
        tot += std::fabs(hPb[i]);
    }
    // UEqn.relax(): the one-pass form equals deviceRelaxDiag on cmptMaxMag3/cmptMin3, and each source moved
    // by delta*U; relaxField's midpoint
    DeviceBuffer<scalar> Us[3] = {DeviceBuffer<scalar>(hU), DeviceBuffer<scalar>(hU), DeviceBuffer<scalar>(hU)};
    DeviceBuffer<scalar> srcs[3] = {DeviceBuffer<scalar>(src), DeviceBuffer<scalar>(src), DeviceBuffer<scalar>(src)};
    DeviceBuffer<scalar> rD, rDelta, rD2, rDelta2, iCmag, iCmin;
    deviceRelaxMomentum(M, dm, iCs, 0.7, Us, srcs, rD, rDelta);
    deviceCmptMaxMag3(iCs[0], iCs[1], iCs[2], iCmag);
    deviceCmptMin3(iCs[0], iCs[1], iCs[2], iCmin);
    deviceRelaxDiag(M, dm, iCs[0], 0.7, rD2, rDelta2, nullptr, iCmag.data(), iCmin.data());
    const std::vector<scalar> hRD = rD.download(), hRD2 = rD2.download(), hDl = rDelta.download(),
                              hS1 = srcs[2].download(), hMag = iCmag.download(), hMin = iCmin.download();
    double relaxErr = (hMag[0] == 0.7 && hMin[0] == 0.3) ? 0.0 : 1.0;
    for (int c = 0; c < nC; ++c)
    {
        relaxErr = std::max(relaxErr, std::fabs(hRD[c] - hRD2[c]));
        relaxErr = std::max(relaxErr, std::fabs(hS1[c] - (src[c] + hDl[c] * hU[c])) / (std::fabs(src[c]) + 1.0));
    }
    DeviceBuffer<scalar> pNew(std::vector<scalar>(nC, 3.0)), pPrev(std::vector<scalar>(nC, 1.0));
    relaxField(pNew, pPrev, 0.5);
    for (const scalar v : pNew.download()) relaxErr = std::max(relaxErr, std::fabs(v - 2.0));

    // the hierarchy written and read back (buildOrLoadAMG) and a clone of it (cloneAMG): each solves the same
    // system to the same bits as the hierarchy built above
    const std::vector<scalar> psiRef = psi.download();
    char dirT[] = "/tmp/brae_amgcache_XXXXXX";
    const std::string cacheDir = mkdtemp(dirT) ? std::string(dirT) : std::string("/tmp");
    AMGCacheRead how1 = AMGCacheRead::loaded, how2 = AMGCacheRead::absent;
    bool cacheOk = true;
    {
        AMGData built = buildOrLoadAMG(own, box.neighbour, magSf, nC, cacheDir, true, nullptr, &how1);
        AMGData loaded = buildOrLoadAMG(own, box.neighbour, magSf, nC, cacheDir, false, nullptr, &how2);
        AMGData cl = cloneAMG(loaded);
        for (AMGData* h : {&built, &loaded, &cl})
        {
            amgGalerkin(*h, diagC, upper, upper);
            DeviceBuffer<scalar> x(std::vector<scalar>(nC, 0.0));
            const DeviceSolverPerf q = deviceAMGPCG(A, *h, b, x, static_cast<const scalar*>(nullptr), 1e-8, 0.0, 500,
                                                    false, 4, true, 0);
            cacheOk = cacheOk && q.nIterations == perf.nIterations && x.download() == psiRef;
        }
        cacheOk = cacheOk && how1 == AMGCacheRead::absent && how2 == AMGCacheRead::loaded && cl.nLevels() == amg.nLevels();
    }
    std::remove((cacheDir + "/.mfv_amgcache").c_str());
    rmdir(cacheDir.c_str());

    // an fvSolution GAMG entry read by readGamgControls drives the same solve as the controls set by hand
    const FoamDict fvSolution = parseFoamDict(
        "solvers { p { solver GAMG; smoother GaussSeidel; tolerance 1e-8; relTol 0; nCellsInCoarsestLevel 10; } }");
    const FoamDict& pEntry = *fvSolution.subDict("solvers")->subDict("p");
    const FoamGamgControls fc = readGamgControls(pEntry, pEntry.scalarOr("tolerance", 1e-6), pEntry.scalarOr("relTol", 0.0),
                                                 pEntry.intOr("maxIter", 1000), "p: ");
    GamgControls gcRead;
    gcRead.smoother = fc.smoother;
    gcRead.tolerance = fc.tolerance;
    gcRead.relTol = fc.relTol;
    gcRead.maxIter = fc.maxIter;
    gcRead.minIter = fc.minIter;
    gcRead.nPreSweeps = fc.nPreSweeps;
    gcRead.preSweepsLevelMultiplier = fc.preSweepsLevelMultiplier;
    gcRead.maxPreSweeps = fc.maxPreSweeps;
    gcRead.nPostSweeps = fc.nPostSweeps;
    gcRead.postSweepsLevelMultiplier = fc.postSweepsLevelMultiplier;
    gcRead.maxPostSweeps = fc.maxPostSweeps;
    gcRead.nFinestSweeps = fc.nFinestSweeps;
    gcRead.scaleCorrection = fc.scaleCorrection;
    gcRead.nCellsInCoarsestLevel = fc.nCellsInCoarsestLevel;
    DeviceBuffer<scalar> psiR(std::vector<scalar>(nC, 0.0));
    const SolverPerformance pr = gamgSolveFolded(gamg, diagC, upper, b, psiR, gcRead);
    const bool readOk = pr.nIterations == pg.nIterations && psiR.download() == psiG.download();

    const bool predictorOk = hErr <= 1e-12 * hScale && hbErr == 0.0 && fluxErr <= 1e-14 &&
                             std::fabs(net) <= 1e-12 * tot && !closedVolume && relaxErr <= 1e-15;

    const bool ok = perf.nIterations > 0 && perf.nIterations < 500 && perf.finalResidual < 1e-8 && res < 2e-8 &&
                    predictorOk &&
                    refusesSA && refusesAsym && flux.size() == (size_t)nI && pg.converged && pg.finalResidual < 1e-8 &&
                    pp.converged && pp.finalResidual < 1e-8 && refusesSmoother && cacheOk && readOk;
    std::printf("{\"cells\": %d, \"levels\": %d, \"iterations\": %d, \"initial\": %.6e, \"final\": %.6e, "
                "\"recomputed\": %.6e, \"refuses_sa\": %s, \"refuses_asym\": %s, \"gamg_levels\": %d, "
                "\"gamg_cycles\": %d, \"gamg_final\": %.6e, \"pcg_gamg_iterations\": %d, \"pcg_gamg_final\": %.6e, "
                "\"refuses_smoother\": %s, \"predictor_h_err\": %.3e, \"predictor_flux_err\": %.3e, "
                "\"predictor_net_flux\": %.3e, \"relax_err\": %.3e, \"cache_clone_same_bits\": %s, "
                "\"fvsolution_gamg_same_bits\": %s, \"ok\": %s}\n",
                nC, amg.nLevels(), perf.nIterations, perf.initialResidual, perf.finalResidual, res,
                refusesSA ? "true" : "false", refusesAsym ? "true" : "false", gamg.nLevels(), pg.nIterations,
                pg.finalResidual, pp.nIterations, pp.finalResidual, refusesSmoother ? "true" : "false",
                hErr / (hScale > 0.0 ? hScale : 1.0), fluxErr, net, relaxErr, cacheOk ? "true" : "false",
                readOk ? "true" : "false", ok ? "true" : "false");
    return ok ? 0 : 1;
}
