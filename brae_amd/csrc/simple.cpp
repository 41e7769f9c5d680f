// One SIMPLE iteration on the device, in the order of the reference's host driver cpu::simpleStep
// (src/applications/solvers/simpleFoam/simpleFoam_cpp.cu:15-237) for uncoupled fixedValue/zeroGradient/empty
// patches without MRF, fvOptions or SIMPLEC:
//
//   UEqn        fvm::div(phi, U) upwind [+ linearUpwind's deferred correction] - fvm::laplacian(nuEff, U)
//               [+ its non-orthogonal correction] + the explicit stress -div(nuEff dev2(T(grad U))), relaxed
//               (UEqn_cpp.cu:156-264, linearViscousStress_cpp.cu:105-184, fv_matrix_ops.cuh:161-209);
//   predictor   solve(UEqn == -grad(p)) on a copy of the source, component by component (PBiCGStab);
//   pressure    rAU, HbyA = rAU H(), constrainHbyA, phiHbyA, adjustPhi when p needs a reference
//               (pEqn_cpp.cu:31-198); then nNonOrthogonalCorrectors + 1 times laplacian(rAUf, p) ==
//               div(phiHbyA) with its correction, setReference, fold and the AMG-PCG solve; on the last
//               pass phi = phiHbyA - pEqn.flux() (pEqn_cpp.cu:201-275);
//   correctors  p relaxed, U = HbyA - rAU grad(p), boundaries evaluated (simpleFoam_cpp.cu:176-184).
//
// Every step is a launch on the context's stream; the host reads back only the solvers' stop flags and
// the adjustPhi decision. The assembly kernels are the host loops' arithmetic, so with sequential
// reductions and the FP64 cycle the step equals the oracle's composition of the same solvers to the bit.
#include "simple.hpp"

#include <cstring>

namespace mfv {

namespace {

template <typename T>
void upload(DBuf<T>& d, const std::vector<T>& h, hipStream_t s)
{
    d.upload(h, s);
}

void sizeAll(std::initializer_list<DBuf<double>*> bufs, size_t n)
{
    for (DBuf<double>* b : bufs) b->resize(n ? n : 1);
}

} // namespace

SimpleSolver::SimpleSolver(Ctx& ctx, Mesh& mesh, const Geometry& geom, const std::vector<int>& hUType,
                           const std::vector<double>& hUValue, const std::vector<int>& hPType,
                           const std::vector<double>& hPValue, const SimpleControls& controls, int mergeLevels)
    : c(ctx), m(mesh), g(geom), ctl(controls)
{
    const int n = m.nCells, nF = m.nFaces, nB = m.nBFaces;
    if ((int)hUType.size() != nB || (int)hPType.size() != nB || (int)hUValue.size() != 3 * nB ||
        (int)hPValue.size() != nB)
        invalid("simple: boundary condition arrays must cover every boundary face");
    for (int i = 0; i < nB; ++i)
    {
        for (int t : {hUType[i], hPType[i]})
            if (t != kBcZeroGradient && t != kBcFixedValue && t != kBcEmpty)
                throw Error(kErrUnsupported, "mfv: simple: boundary condition " + std::to_string(t) + " on face " +
                                                 std::to_string(i) +
                                                 " is not carried (zeroGradient, fixedValue/noSlip, empty)");
        if ((hUType[i] == kBcEmpty) != (hPType[i] == kBcEmpty))
            invalid("simple: U and p must agree on the empty faces");
    }
    if (ctl.pRefCell >= n) invalid("simple: pRefCell outside the mesh");
    if (ctl.nNonOrthCorr < 0 || ctl.nNonOrthCorr > 7) invalid("simple: nNonOrthogonalCorrectors must be 0..7");
    if (!g.Sf || !g.Cf || !g.magSf || !g.C || !g.V || !g.w || !g.dc || !g.bDelta || (ctl.correctedLaplacian &&
                                                                                     (!g.ndc || !g.corrVec)))
        invalid("simple: geometry array missing");
    std::vector<int> gT(nB), pgT(nB), emp(nB), take(nB), adj(nB);
    for (int i = 0; i < nB; ++i)
    {
        emp[i] = hUType[i] == kBcEmpty;
        gT[i] = emp[i] ? kBcEmpty : kBcFixedValue;   // the gradients take every non-empty patch value
        pgT[i] = gT[i];
        take[i] = hUType[i] == kBcFixedValue;         // constrainHbyA: U is not assignable there
        adj[i] = hUType[i] != kBcFixedValue;          // adjustPhi: outflow that may be scaled
    }
    upload(uType, hUType, c.stream);
    upload(pType, hPType, c.stream);
    upload(gType, gT, c.stream);
    upload(pgType, pgT, c.stream);
    upload(isEmpty, emp, c.stream);
    upload(takeU, take, c.stream);
    upload(adjustable, adj, c.stream);
    upload(uValue, hUValue, c.stream);
    upload(pValue, hPValue, c.stream);

    sizeAll({&D, &lD, &nuC, &nuCneg, &pPrev, &Dm, &rAU, &pD, &psrc, &diagC, &b}, n);
    sizeAll({&Up, &Lo, &lU, &nuF, &nuFneg, &phiHbyA, &rAUf, &pU, &ffc}, nF);
    sizeAll({&nuB, &nuBF, &icMag, &icMin, &phiHbyAB, &rAUb, &piC, &pbC}, nB);
    for (int k = 0; k < 3; ++k)
    {
        sizeAll({&iC[k], &bC[k], &HbyAb[k], &bd[k]}, nB);
        sizeAll({&src[k], &srcP[k], &dv[k], &corr[k], &HbyA[k]}, n);
        g3[k].resize(3 * (size_t)std::max(n, 1));
    }
    gradU.resize(9 * (size_t)std::max(n, 1));
    sigB.resize(9 * (size_t)std::max(nB, 1));
    gradP.resize(3 * (size_t)std::max(n, 1));
    decision.resize(4);

    // the pressure hierarchy on the mesh's face areas (brae's pairing), kept for every iteration
    std::vector<double> w(nF);
    if (nF) hipCheck(hipMemcpyAsync(w.data(), g.magSf, sizeof(double) * nF, hipMemcpyDeviceToHost, c.stream), "weights");
    hipCheck(hipStreamSynchronize(c.stream), "simple setup");
    AmgBuildControls bc;
    if (ctl.pSolver == PSolver::GamgRef)
    {
        // the host CPU path's GAMG: gamg.cu's agglomeratePair on |Sf| down to 10 cells, single passes
        bc.kind = AggKind::Pair;
        bc.target = 10;
        bc.mergeLevels = 1;
    }
    else
    {
        bc.kind = AggKind::Brae;
        bc.target = 64;
        bc.mergeLevels = std::max(1, mergeLevels);
    }
    amgBuild(c, m, amg, w, bc);
}

SimpleResult SimpleSolver::step(const SimpleFields& f, const double* nuEff, const double* nuEffB)
{
    const int n = m.nCells, nF = m.nFaces, nB = m.nBFaces;
    for (int k = 0; k < 3; ++k)
        if (!f.U[k] || (nB && !f.Ub[k])) invalid("simple step: velocity field missing");
    if (!f.p || (nB && !f.pb) || (nF && !f.phi) || (nB && !f.phiB)) invalid("simple step: field missing");
    SimpleResult res;
    const double* bMagSf = g.magSf + nF;
    const double* dcU = ctl.correctedLaplacian ? g.ndc : g.dc;

    // nuEff for this iteration: the caller's (nu + nut of the previous iteration), or the laminar nu
    if (nuEff)
        launchCopy(c, n, nuEff, nuC.p);
    else
        launchFill(c, n, ctl.nu, nuC.p);
    if (nuEffB)
        launchCopy(c, nB, nuEffB, nuB.p);
    else
        launchFill(c, nB, ctl.nu, nuB.p);

    // prevIter of p, and p's patch values at the top of the iteration
    launchCopy(c, n, f.p, pPrev.p);
    {
        const double* cells[1] = {f.p};
        double* outs[1] = {f.pb};
        launchEvaluate(c, m, 1, 1, pType.p, pValue.p, cells, outs);
    }

    // ---- UEqn ----
    ProfScope psU(c, kProfSimpleUEqn);
    double* ic[3] = {iC[0].p, iC[1].p, iC[2].p};
    double* bcf[3] = {bC[0].p, bC[1].p, bC[2].p};
    double* s[3] = {src[0].p, src[1].p, src[2].p};
    const double* U[3] = {f.U[0], f.U[1], f.U[2]};
    const double* Ub[3] = {f.Ub[0], f.Ub[1], f.Ub[2]};
    launchDivUpwind(c, m, f.phi, D.p, Up.p, Lo.p);
    launchMomBoundary(c, nB, uType.p, uValue.p, f.phiB, nuB.p, bMagSf, g.bDelta, ic, bcf);
    {
        double* g3o[3] = {g3[0].p, g3[1].p, g3[2].p};
        launchGradGauss3(c, m, gType.p, Ub, g.Sf, g.w, g.V, U, g3o, gradU.p);
    }
    for (int k = 0; k < 3; ++k) launchZero(c, n, s[k]);
    const double* g3p[3] = {g3[0].p, g3[1].p, g3[2].p};
    if (ctl.linearUpwind && nF)
    {
        double* cr[3] = {corr[0].p, corr[1].p, corr[2].p};
        launchLinearUpwindCorr(c, m, f.phi, g.Cf, g.C, 3, g3p, cr);
        for (int k = 0; k < 3; ++k) launchSubInPlace(c, n, cr[k], s[k]);
    }
    // - fvm::laplacian(nuEff, U): face nuEff interpolated, the coefficients subtracted (addEqual with -1)
    launchInterpolate(c, m, g.w, nuC.p, nuF.p, nullptr);
    if (nF)
    {
        launchLaplacian(c, m, dcU, nuF.p, g.magSf, lD.p, lU.p, lU.p);
        launchAxpy(c, n, -1.0, lD.p, D.p);
        launchAxpy(c, nF, -1.0, lU.p, Up.p);
        launchAxpy(c, nF, -1.0, lU.p, Lo.p);
    }
    if (ctl.correctedLaplacian && nF)
    {
        // its explicit correction enters the momentum source with + (the laplacian is subtracted): with the
        // face viscosity negated the kernel's source -= div(ffc) adds exactly div of the true correction
        launchNegate(c, nF, nuF.p, nuFneg.p);
        for (int k = 0; k < 3; ++k) launchLaplacianCorr(c, m, g3[k].p, g.w, g.corrVec, nuFneg.p, g.magSf, ffc.p, s[k]);
    }
    // the explicit stress: boundary stress from the corrected boundary gradient, then V*div in one pass
    launchStressBoundary(c, m, uType.p, g.Sf, g.magSf, g.bDelta, U, Ub, gradU.p, nuB.p, sigB.p);
    launchTensorDiv(c, m, g.w, g.Sf, isEmpty.p, gradU.p, nuC.p, sigB.p, dv[0].p, dv[1].p, dv[2].p);
    {
        const double* d3[3] = {dv[0].p, dv[1].p, dv[2].p};
        launchStressSource(c, n, g.V, d3, s);
    }
    // UEqn.relax(): into Dm, which then holds the momentum matrix's diagonal
    double* Dmom = D.p;
    if (ctl.relaxU > 0.0 && ctl.relaxU < 1.0)
    {
        launchCmpt3(c, nB, ic[0], ic[1], ic[2], false, icMag.p);
        launchCmpt3(c, nB, ic[0], ic[1], ic[2], true, icMin.p);
        launchRelaxDiag(c, m, Up.p, Lo.p, D.p, nullptr, icMag.p, icMin.p, nullptr, ctl.relaxU, 3, U, s, Dm.p, nullptr);
        Dmom = Dm.p;
    }

    psU.close();

    // ---- momentum predictor: solve(UEqn == -grad(p)) on a copy of the source ----
    if (ctl.momentumPredictor)
    {
        ProfScope ps(c, kProfSimpleUSolve);
        launchGradGauss(c, m, pgType.p, f.pb, g.Sf, g.w, g.V, f.p, gradP.p);
        double* sp[3] = {srcP[0].p, srcP[1].p, srcP[2].p};
        const double* sc[3] = {s[0], s[1], s[2]};
        launchGradPSource(c, n, gradP.p, g.V, sc, sp);
        const LduView Am{n, nF, Dmom, Up.p, Lo.p};
        for (int k = 0; k < 3; ++k)
        {
            const PcgPerf pf = pbicgstabSolve(c, m, Am, sp[k], ic[k], bcf[k], f.U[k], ctl.U);
            res.UInitial[k] = pf.initialResidual;
            res.UFinal[k] = pf.finalResidual;
            res.UIters[k] = pf.nIterations;
        }
        double* ubo[3] = {f.Ub[0], f.Ub[1], f.Ub[2]};
        launchEvaluate(c, m, 3, 3, uType.p, uValue.p, U, ubo);
    }

    // ---- pressure predictor ----
    ProfScope psPred(c, kProfSimplePPredictor);
    launchFoldBoundaryDiag(c, m, ic[0], ic[1], ic[2], Dmom, pD.p);
    launchReciprocalV(c, m, pD.p, g.V, rAU.p);
    double* hb[3] = {HbyA[0].p, HbyA[1].p, HbyA[2].p};
    {
        const double* bdp[3] = {bd[0].p, bd[1].p, bd[2].p};
        const double* bcp[3] = {bcf[0], bcf[1], bcf[2]};
        const double* sc[3] = {s[0], s[1], s[2]};
        for (int k = 0; k < 3; ++k) launchBoundaryDiagTerm(c, nB, ic[0], ic[1], ic[2], k, bd[k].p);
        launchMatrixH3(c, m, Up.p, Lo.p, U, sc, bdp, bcp, g.V, rAU.p, nullptr, hb);
    }
    for (int k = 0; k < 3; ++k) launchConstrainHbyA(c, m, takeU.p, f.Ub[k], HbyA[k].p, HbyAb[k].p);
    if (nF) launchVectorFlux(c, m, g.w, g.Sf, hb[0], hb[1], hb[2], phiHbyA.p);
    if (nB) launchBoundaryFlux(c, m, g.Sf + 3 * (size_t)nF, HbyAb[0].p, HbyAb[1].p, HbyAb[2].p, isEmpty.p, phiHbyAB.p);
    if (ctl.pRefCell >= 0)
    {
        launchAdjustPhi(c, nB, adjustable.p, phiHbyAB.p, nF, phiHbyA.p, decision.p);
        hipCheck(hipMemcpyAsync(c.mailbox, decision.p, 3 * sizeof(double), hipMemcpyDeviceToHost, c.stream),
                 "adjust phi read-back");
        hipCheck(hipStreamSynchronize(c.stream), "adjust phi sync");
        if (c.mailbox[1] != 0.0)
            invalid("simple: adjustPhi: the continuity error cannot be removed by adjusting the outflow "
                    "(adjustPhi.C): check the velocity boundary conditions");
        res.massCorr = c.mailbox[0];
    }

    // ---- the pressure equation, nNonOrthogonalCorrectors + 1 times ----
    launchInterpolate(c, m, g.w, rAU.p, rAUf.p, rAUb.p);
    psPred.close();
    for (int corrPass = 0; corrPass <= ctl.nNonOrthCorr; ++corrPass)
    {
        ProfScope psAsm(c, kProfSimplePAssembly);
        if (nF) launchLaplacian(c, m, dcU, rAUf.p, g.magSf, pD.p, pU.p, pU.p);
        else launchZero(c, n, pD.p);
        launchLaplacianBCoeffs(c, nB, pType.p, pValue.p, rAUb.p, bMagSf, g.bDelta, piC.p, pbC.p);
        launchZero(c, n, psrc.p);
        if (ctl.correctedLaplacian && nF)
        {
            launchGradGauss(c, m, pgType.p, f.pb, g.Sf, g.w, g.V, f.p, gradP.p);
            launchLaplacianCorr(c, m, gradP.p, g.w, g.corrVec, rAUf.p, g.magSf, ffc.p, psrc.p);
        }
        launchDivSource(c, m, phiHbyA.p, phiHbyAB.p, g.V, 1.0, psrc.p);
        if (ctl.pRefCell >= 0) launchSetReference(c, pD.p, psrc.p, ctl.pRefCell, ctl.pRefValue);
        launchFold(c, m, pD.p, psrc.p, piC.p, pbC.p, diagC.p, b.p);
        psAsm.close();
        if (corrPass == 0 && ctl.pSolver == PSolver::AmgPcg)
        {
            ProfScope ps(c, kProfSimpleGalerkin);
            amgGalerkin(c, amg, diagC.p, pU.p, ctl.P.prec);
        }
        ProfScope psSolve(c, kProfSimplePSolve);
        PcgPerf pf;
        if (ctl.pSolver == PSolver::GamgRef)
        {
            // gamg(pEqn, p, ..., nCellsInCoarsest 10, nPreSweeps 0, nPostSweeps 2, nFinestSweeps 2)
            // (simpleFoam_cpp.cu:156-168)
            GamgRefControls gr;
            gr.tolerance = ctl.P.tolerance;
            gr.relTol = ctl.P.relTol;
            gr.maxIter = ctl.P.maxIter;
            gr.minIter = ctl.P.minIter;
            pf = gamgRefSolve(c, m, amg, diagC.p, pU.p, b.p, f.p, gr);
        }
        else
            pf = amgPcgSolve(c, m, amg, diagC.p, pU.p, b.p, f.p, nullptr, ctl.P);
        psSolve.close();
        if (corrPass == 0) res.pInitial = pf.initialResidual;
        res.pFinal = pf.finalResidual;
        res.pIters[corrPass] = pf.nIterations;
        {
            const double* cells[1] = {f.p};
            double* outs[1] = {f.pb};
            launchEvaluate(c, m, 1, 1, pType.p, pValue.p, cells, outs);
        }
        if (corrPass == ctl.nNonOrthCorr)
        {
            ProfScope ps(c, kProfSimpleCorrect);
            if (nF)
                launchFluxCorrect(c, m, LduView{n, nF, diagC.p, pU.p, pU.p}, f.p,
                                  ctl.correctedLaplacian ? ffc.p : nullptr, phiHbyA.p, 1.0, f.phi);
            launchBoundaryFluxCorr(c, m, phiHbyAB.p, piC.p, pbC.p, f.p, f.phiB);
        }
    }

    // ---- p.relax(), then U = HbyA - rAU grad(p) ----
    ProfScope psCorr(c, kProfSimpleCorrect);
    if (ctl.relaxP > 0.0 && ctl.relaxP < 1.0) launchRelaxField(c, n, f.p, pPrev.p, ctl.relaxP);
    {
        const double* cells[1] = {f.p};
        double* outs[1] = {f.pb};
        launchEvaluate(c, m, 1, 1, pType.p, pValue.p, cells, outs);
    }
    launchGradGauss(c, m, pgType.p, f.pb, g.Sf, g.w, g.V, f.p, gradP.p);
    double* Uo[3] = {f.U[0], f.U[1], f.U[2]};
    const double* hbc[3] = {hb[0], hb[1], hb[2]};
    launchCorrectU(c, n, rAU.p, gradP.p, hbc, Uo);
    double* ubo[3] = {f.Ub[0], f.Ub[1], f.Ub[2]};
    launchEvaluate(c, m, 3, 3, uType.p, uValue.p, U, ubo);
    psCorr.close();
    hipCheck(hipStreamSynchronize(c.stream), "simple step");
    return res;
}

} // namespace mfv
