// The standard k-epsilon model on the device, as the reference's host SIMPLE driver calls it after each
// iteration: kepsilon::correct (src/TurbulenceModels/turbulenceModels/RAS/kEpsilon/k_epsilon.cuh:221-337) with
// the epsilonWallFunction's near-wall epsilon and production, the nutkWallFunction
// (src/TurbulenceModels/turbulenceModels/derivedFvPatchFields/wallFunctions/nut_wall_function.cuh:31-35) and
// fvMatrix::setValues (src/finiteVolume/fvMatrices/fv_matrix_ops.cuh:110-148); and the near-wall distance
// (src/finiteVolume/fvMesh/near_wall_dist.cuh:31-177) on the host, once per mesh.
//
// Incompressible, upwind convection of k and epsilon, uncoupled fixedValue/zeroGradient/empty patches (the
// wall functions' patches carry zeroGradient values), wall functions on the faces of `wall` patches.
// Every per-cell sum is a gather in the host loop's face order and the library is built without FMA
// contraction, so the model's arithmetic is the host's term for term; only the wall functions' pow and log
// come from the device's math library, whose last bit may differ from the host's.
#include "reduce.hpp"
#include "simple.hpp"

#include <cmath>

namespace mfv {

namespace {

constexpr int kTpb = 256;
inline int blocks(int n) { return (n + kTpb - 1) / kTpb; }

// GbyNu = gradU && devTwoSymm(gradU) (k_epsilon.cuh:62-76), G = nut*GbyNu, and divU = fvc::div(phi)
// (fvc.cu:706-733: internal faces in face order, then the non-empty boundary faces, over V)
__global__ __launch_bounds__(kTpb) void keProductionK(int n, const label* __restrict__ ownerStart,
                                                      const label* __restrict__ losortStart,
                                                      const label* __restrict__ losort,
                                                      const label* __restrict__ bndStart,
                                                      const label* __restrict__ bndFaces,
                                                      const int* __restrict__ isEmpty,
                                                      const double* __restrict__ gradU, const double* __restrict__ nut,
                                                      const double* __restrict__ phi, const double* __restrict__ phiB,
                                                      const double* __restrict__ V, double* __restrict__ gbn,
                                                      double* __restrict__ G, double* __restrict__ divU)
{
    const int c = xcdBlock() * kTpb + threadIdx.x;
    if (c >= n) return;
    double t[9];
    for (int q = 0; q < 9; ++q) t[q] = gradU[(size_t)q * n + c];
    const double t23 = (2.0 / 3.0) * (t[0] + t[4] + t[8]);
    double d[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) d[i * 3 + j] = t[i * 3 + j] + t[j * 3 + i];
    d[0] -= t23;
    d[4] -= t23;
    d[8] -= t23;
    double s = t[0] * d[0];
    for (int q = 1; q < 9; ++q) s += t[q] * d[q];
    gbn[c] = s;
    G[c] = nut[c] * s;
    double dv = 0.0;
    for (int k = losortStart[c]; k < losortStart[c + 1]; ++k) dv -= phi[losort[k]];
    for (int f = ownerStart[c]; f < ownerStart[c + 1]; ++f) dv += phi[f];
    for (int k = bndStart[c]; k < bndStart[c + 1]; ++k)
    {
        const int i = bndFaces[k];
        if (!isEmpty[i]) dv += phiB[i];
    }
    divU[c] = dv / V[c];
}

struct WallArgs
{
    double Cmu25, Cmu75, kappa, E, yplLam, nu;
};

// nutkWallFunction's value on one face (nut_wall_function.cuh:31-35)
__device__ __forceinline__ double nutkWall(const WallArgs& a, double y, double kc)
{
    const double yPlus = a.Cmu25 * y * sqrt(kc) / a.nu;
    return (yPlus > a.yplLam) ? (a.nu * yPlus * a.kappa / log(fmax(a.E * yPlus, 1.0 + 1e-4)) - a.nu) : 0.0;
}

// the epsilonWallFunction (k_epsilon.cuh:242-274): per cell with wall faces, the corner-weighted near-wall
// epsilon and production over its wall faces ascending, written over epsilon and G; the cell is marked
__global__ __launch_bounds__(kTpb) void keWallK(int n, const label* __restrict__ bndStart,
                                                const label* __restrict__ bndFaces, const int* __restrict__ isWall,
                                                const double* __restrict__ y, const double* __restrict__ bDelta,
                                                const double* __restrict__ Ux, const double* __restrict__ Uy,
                                                const double* __restrict__ Uz, const double* __restrict__ Ubx,
                                                const double* __restrict__ Uby, const double* __restrict__ Ubz,
                                                const double* __restrict__ k, WallArgs a, double* __restrict__ G,
                                                double* __restrict__ eps, int* __restrict__ wallCell)
{
    const int c = blockIdx.x * kTpb + threadIdx.x;
    if (c >= n) return;
    const int b0 = bndStart[c], b1 = bndStart[c + 1];
    int nw = 0;
    for (int q = b0; q < b1; ++q) nw += isWall[bndFaces[q]] != 0;
    if (nw == 0)
    {
        wallCell[c] = 0;
        return;
    }
    const double w = 1.0 / nw, kc = k[c];
    const double ux = Ux[c], uy = Uy[c], uz = Uz[c];
    double e0 = 0.0, g0 = 0.0;
    for (int q = b0; q < b1; ++q)
    {
        const int i = bndFaces[q];
        if (!isWall[i]) continue;
        const double yi = y[i];
        const double nutw = nutkWall(a, yi, kc);
        const double dc = bDelta[i];
        const double gx = (Ubx[i] - ux) * dc, gy = (Uby[i] - uy) * dc, gz = (Ubz[i] - uz) * dc;
        const double magGradUw = sqrt(gx * gx + gy * gy + gz * gz);
        e0 += w * a.Cmu75 * pow(kc, 1.5) / (a.kappa * yi);
        g0 += w * (nutw + a.nu) * magGradUw * a.Cmu25 * sqrt(kc) / (a.kappa * yi);
    }
    G[c] = g0;
    eps[c] = e0;
    wallCell[c] = 1;
}

// the face diffusivity of an equation: interpolate(nut/sigma + nu) (k_epsilon.cuh:100-134)
__global__ __launch_bounds__(kTpb) void keFaceDiffK(int nF, const label* __restrict__ own,
                                                    const label* __restrict__ nei, const double* __restrict__ w,
                                                    const double* __restrict__ nut, double sigma, double nu,
                                                    double* __restrict__ Df)
{
    const int f = blockIdx.x * kTpb + threadIdx.x;
    if (f >= nF) return;
    const double wf = w[f];
    Df[f] = wf * (nut[own[f]] / sigma + nu) + (1.0 - wf) * (nut[nei[f]] / sigma + nu);
}

// an equation's boundary coefficients: fvm::div's (fixedValue: 0, value; zeroGradient: 1, 0) plus -1 times
// fvm::laplacian's with the patch diffusivity nut_b/sigma + nu (fixedValue: -deltaCoeffs,
// deltaCoeffs*value); empty faces zero
__global__ __launch_bounds__(kTpb) void keBoundaryK(int nB, const int* __restrict__ type,
                                                    const double* __restrict__ value, const double* __restrict__ phib,
                                                    const double* __restrict__ nutb, double sigma, double nu,
                                                    const double* __restrict__ bMagSf,
                                                    const double* __restrict__ bDelta, double* __restrict__ iC,
                                                    double* __restrict__ bC)
{
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= nB) return;
    const int t = type[i];
    if (t == kBcEmpty)
    {
        iC[i] = 0.0;
        bC[i] = 0.0;
        return;
    }
    const double ph = phib[i];
    const double pGamma = (nutb[i] / sigma + nu) * bMagSf[i];
    double vIC = 1.0, vBC = 0.0, gIC = 0.0, gBC = 0.0;
    if (t == kBcFixedValue)
    {
        const double v = value[i];
        vIC = 0.0;
        vBC = v;
        gIC = -bDelta[i];
        gBC = bDelta[i] * v;
    }
    const double lIC = pGamma * gIC;
    const double lBC = (-pGamma) * gBC;
    iC[i] = ph * vIC + -1.0 * lIC;
    bC[i] = (-ph) * vBC + -1.0 * lBC;
}

// the implicit and explicit source terms per cell: epsilon (k_epsilon.cuh:182-186 and 288-293)
//   D += (C2 eps/k) V; S = (C1 Cmu k GbyNu) V; D += V max(sp, 0); S -= V min(sp, 0) eps, sp = ((2/3)C1 - C3) divU
// or k (k_epsilon.cuh:155-159 and 303-308)
//   D += (eps/k) V; S = G V; D += V max(sp, 0); S -= V min(sp, 0) k, sp = (2/3) divU
template <bool EPS>
__global__ __launch_bounds__(kTpb) void keSourceK(int n, double C1, double C2, double Cmu, double spCoeff,
                                                  const double* __restrict__ V, const double* __restrict__ eps,
                                                  const double* __restrict__ k, const double* __restrict__ prod,
                                                  const double* __restrict__ divU, double* __restrict__ D,
                                                  double* __restrict__ S)
{
    const int c = blockIdx.x * kTpb + threadIdx.x;
    if (c >= n) return;
    const double v = V[c], e = eps[c], kc = k[c];
    double d = D[c], s = 0.0;
    if (EPS)
    {
        d += (C2 * e / kc) * v;
        s += (C1 * Cmu * kc * prod[c]) * v;
    }
    else
    {
        d += (e / kc) * v;
        s += prod[c] * v;
    }
    const double sp = spCoeff * divU[c];
    d += v * fmax(sp, 0.0);
    s -= v * fmin(sp, 0.0) * (EPS ? e : kc);
    D[c] = d;
    S[c] = s;
}

// fvMatrix::setValues, first half (fv_matrix_ops.cuh:120-141): a cell next to constrained cells takes their
// couplings into its source, in face order (the order the host's constrained cells ascending reach it)
__global__ __launch_bounds__(kTpb) void keSetValuesSrcK(int n, const label* __restrict__ ownerStart,
                                                        const label* __restrict__ losortStart,
                                                        const label* __restrict__ losort,
                                                        const label* __restrict__ losortOwner,
                                                        const label* __restrict__ neighbour,
                                                        const double* __restrict__ upper,
                                                        const double* __restrict__ lower,
                                                        const int* __restrict__ wallCell,
                                                        const double* __restrict__ value, double* __restrict__ S)
{
    const int c = blockIdx.x * kTpb + threadIdx.x;
    if (c >= n || wallCell[c]) return;
    double s = S[c];
    bool any = false;
    for (int k = losortStart[c]; k < losortStart[c + 1]; ++k)
    {
        const int o = losortOwner[k];
        if (wallCell[o])
        {
            s -= lower[losort[k]] * value[o];
            any = true;
        }
    }
    for (int f = ownerStart[c]; f < ownerStart[c + 1]; ++f)
    {
        const int o = neighbour[f];
        if (wallCell[o])
        {
            s -= upper[f] * value[o];
            any = true;
        }
    }
    if (any) S[c] = s;
}

// setValues, second half: the constrained cells' faces lose their coefficients, their boundary faces theirs,
// and psi = value, source = value*diag (fv_matrix_ops.cuh:136-147)
__global__ __launch_bounds__(kTpb) void keSetValuesFaceK(int nF, const label* __restrict__ own,
                                                         const label* __restrict__ nei,
                                                         const int* __restrict__ wallCell, double* __restrict__ upper,
                                                         double* __restrict__ lower)
{
    const int f = blockIdx.x * kTpb + threadIdx.x;
    if (f >= nF) return;
    if (wallCell[own[f]] || wallCell[nei[f]])
    {
        upper[f] = 0.0;
        lower[f] = 0.0;
    }
}

__global__ __launch_bounds__(kTpb) void keSetValuesBndK(int nB, const label* __restrict__ bfc,
                                                        const int* __restrict__ wallCell, double* __restrict__ iC,
                                                        double* __restrict__ bC)
{
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= nB) return;
    if (wallCell[bfc[i]])
    {
        iC[i] = 0.0;
        bC[i] = 0.0;
    }
}

__global__ __launch_bounds__(kTpb) void keSetValuesCellK(int n, const int* __restrict__ wallCell,
                                                         const double* __restrict__ value,
                                                         const double* __restrict__ D, double* __restrict__ S,
                                                         double* __restrict__ psi)
{
    const int c = blockIdx.x * kTpb + threadIdx.x;
    if (c >= n || !wallCell[c]) return;
    const double v = value[c];
    psi[c] = v;
    S[c] = v * D[c];
}

// Foam::bound as the host driver applies it: max(x, 1e-15)
__global__ __launch_bounds__(kTpb) void keBoundK(int n, double* __restrict__ x)
{
    const int c = blockIdx.x * kTpb + threadIdx.x;
    if (c < n) x[c] = fmax(x[c], 1e-15);
}

// correctNut: Cmu k^2/epsilon in the cells (k_epsilon.cuh:78-87)
__global__ __launch_bounds__(kTpb) void keNutK(int n, double Cmu, const double* __restrict__ k,
                                               const double* __restrict__ eps, double* __restrict__ nut)
{
    const int c = blockIdx.x * kTpb + threadIdx.x;
    if (c < n) nut[c] = Cmu * k[c] * k[c] / eps[c];
}

// ... and the nutkWallFunction of the new k on the wall faces (k_epsilon.cuh:330-335)
__global__ __launch_bounds__(kTpb) void keNutWallK(int nB, const label* __restrict__ bfc,
                                                   const int* __restrict__ isWall, const double* __restrict__ y,
                                                   const double* __restrict__ k, WallArgs a,
                                                   double* __restrict__ nutb)
{
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= nB || !isWall[i]) return;
    nutb[i] = nutkWall(a, y[i], k[bfc[i]]);
}

// nuEff = nu + nut, cells and boundary faces (simpleFoam_cpp.cu:27-46)
__global__ __launch_bounds__(kTpb) void keNuEffK(int n, double nu, const double* __restrict__ nut,
                                                 double* __restrict__ out)
{
    const int c = blockIdx.x * kTpb + threadIdx.x;
    if (c < n) out[c] = nu + nut[c];
}

double yPlusLamHost(double kappa, double E)
{
    double ypl = 11.0;
    for (int i = 0; i < 10; ++i) ypl = std::log(std::fmax(E * ypl, 1.0)) / kappa;
    return ypl;
}

} // namespace

KEpsilon::KEpsilon(Ctx& ctx, Mesh& mesh, const Geometry& geom, const std::vector<int>& hIsWall,
                   const std::vector<double>& hY, const std::vector<int>& hKType, const std::vector<double>& hKValue,
                   const std::vector<int>& hEpsType, const std::vector<double>& hEpsValue,
                   const KEpsilonControls& controls)
    : c(ctx), m(mesh), g(geom), ctl(controls)
{
    const int n = m.nCells, nF = m.nFaces, nB = m.nBFaces;
    if ((int)hIsWall.size() != nB || (int)hY.size() != nB || (int)hKType.size() != nB ||
        (int)hKValue.size() != nB || (int)hEpsType.size() != nB || (int)hEpsValue.size() != nB)
        invalid("kepsilon: boundary arrays must cover every boundary face");
    for (int i = 0; i < nB; ++i)
    {
        for (int t : {hKType[i], hEpsType[i]})
            if (t != kBcZeroGradient && t != kBcFixedValue && t != kBcEmpty)
                throw Error(kErrUnsupported, "mfv: kepsilon: boundary condition " + std::to_string(t) + " on face " +
                                                 std::to_string(i) +
                                                 " is not carried (zeroGradient, fixedValue, empty)");
        if ((hKType[i] == kBcEmpty) != (hEpsType[i] == kBcEmpty))
            invalid("kepsilon: k and epsilon must agree on the empty faces");
        if (hIsWall[i] && (hKType[i] != kBcZeroGradient || hEpsType[i] != kBcZeroGradient))
            invalid("kepsilon: a wall face carries the wall functions: k and epsilon zeroGradient there");
        if (hIsWall[i] && !(hY[i] > 0.0)) invalid("kepsilon: a wall face needs a positive near-wall distance");
    }
    if (!(ctl.nu > 0.0)) invalid("kepsilon: nu must be positive");
    if (!g.Sf || !g.magSf || !g.V || !g.w || !g.dc || !g.bDelta) invalid("kepsilon: geometry array missing");
    std::vector<int> emp(nB);
    for (int i = 0; i < nB; ++i) emp[i] = hKType[i] == kBcEmpty;
    isWall.upload(hIsWall, c.stream);
    yWall.upload(hY, c.stream);
    kType.upload(hKType, c.stream);
    epsType.upload(hEpsType, c.stream);
    kValue.upload(hKValue, c.stream);
    epsValue.upload(hEpsValue, c.stream);
    isEmpty.upload(emp, c.stream);
    std::vector<int> gT(nB);
    for (int i = 0; i < nB; ++i) gT[i] = emp[i] ? kBcEmpty : kBcFixedValue;
    gType.upload(gT, c.stream);
    const size_t N = std::max(n, 1), F = std::max(nF, 1), B = std::max(nB, 1);
    for (DBuf<double>* b : {&gbn, &G, &divU, &D, &Dr, &S, &lD}) b->resize(N);
    for (DBuf<double>* b : {&Up, &Lo, &Df, &lU}) b->resize(F);
    for (DBuf<double>* b : {&iC, &bC, &nutB0}) b->resize(B);
    wallCell.resize(N);
    // the cells the wall functions constrain: those with a wall face
    std::vector<char> wc(n, 0);
    for (int i = 0; i < nB; ++i)
        if (hIsWall[i]) wc[m.hBFaceCell[i]] = 1;
    nWallCells = 0;
    for (int cc = 0; cc < n; ++cc) nWallCells += wc[cc];
    hipCheck(hipStreamSynchronize(c.stream), "kepsilon setup");
}

void KEpsilon::nuEff(const double* nut, const double* nutb, double* nuEffC, double* nuEffB)
{
    if (m.nCells) keNuEffK<<<blocks(m.nCells), kTpb, 0, c.stream>>>(m.nCells, ctl.nu, nut, nuEffC);
    if (m.nBFaces) keNuEffK<<<blocks(m.nBFaces), kTpb, 0, c.stream>>>(m.nBFaces, ctl.nu, nutb, nuEffB);
    hipCheck(hipGetLastError(), "nuEff");
}

// one scalar equation: assemble, relax, constrain (epsilon), solve, bound, evaluate its patches
PcgPerf KEpsilon::solveScalar(bool isEps, const double* phi, const double* phiB, const double* nut, double* psi,
                              double* psib, const double* prod, const double* eps, const double* k)
{
    const int n = m.nCells, nF = m.nFaces, nB = m.nBFaces;
    const KEpsilonCoeffs& co = ctl.co;
    const double sigma = isEps ? co.sigmaEps : co.sigmaK;
    const int* type = isEps ? epsType.p : kType.p;
    const double* value = isEps ? epsValue.p : kValue.p;
    // div(phi, psi) - laplacian(D, psi): upwind, then the orthogonal laplacian subtracted (addEqual with -1)
    launchDivUpwind(c, m, phi, D.p, Up.p, Lo.p);
    if (nB)
        keBoundaryK<<<blocks(nB), kTpb, 0, c.stream>>>(nB, type, value, phiB, nutB0.p, sigma, ctl.nu,
                                                       g.magSf + nF, g.bDelta, iC.p, bC.p);
    if (nF)
    {
        keFaceDiffK<<<blocks(nF), kTpb, 0, c.stream>>>(nF, m.owner.p, m.neighbour.p, g.w, nut, sigma, ctl.nu, Df.p);
        launchLaplacian(c, m, g.dc, Df.p, g.magSf, lD.p, lU.p, lU.p);
        launchAxpy(c, n, -1.0, lD.p, D.p);
        launchAxpy(c, nF, -1.0, lU.p, Up.p);
        launchAxpy(c, nF, -1.0, lU.p, Lo.p);
    }
    if (n)
    {
        if (isEps)
            keSourceK<true><<<blocks(n), kTpb, 0, c.stream>>>(n, co.C1, co.C2, co.Cmu, (2.0 / 3.0) * co.C1 - co.C3,
                                                              g.V, eps, k, prod, divU.p, D.p, S.p);
        else
            keSourceK<false><<<blocks(n), kTpb, 0, c.stream>>>(n, co.C1, co.C2, co.Cmu, 2.0 / 3.0, g.V, eps, k, prod,
                                                               divU.p, D.p, S.p);
    }
    hipCheck(hipGetLastError(), "kepsilon assembly");
    // relax (fvMatrix::relax: a no-op for alpha <= 0), cmptMagMax/cmptMin of a scalar being |iC| and iC
    const double alpha = isEps ? ctl.relaxEps : ctl.relaxK;
    const double* Dm = D.p;
    if (alpha > 0.0)
    {
        double* src[1] = {S.p};
        const double* ps[1] = {psi};
        launchRelaxDiag(c, m, Up.p, Lo.p, D.p, iC.p, nullptr, nullptr, nullptr, alpha, 1, ps, src, Dr.p, nullptr);
        Dm = Dr.p;
    }
    if (isEps && nWallCells)
    {
        // setValues at the near-wall cells: the value is epsilon there (the wall function just wrote it)
        keSetValuesSrcK<<<blocks(n), kTpb, 0, c.stream>>>(n, m.ownerStart.p, m.losortStart.p, m.losort.p,
                                                          m.losortOwner.p, m.neighbour.p, Up.p, Lo.p, wallCell.p,
                                                          psi, S.p);
        if (nF)
            keSetValuesFaceK<<<blocks(nF), kTpb, 0, c.stream>>>(nF, m.owner.p, m.neighbour.p, wallCell.p, Up.p, Lo.p);
        if (nB) keSetValuesBndK<<<blocks(nB), kTpb, 0, c.stream>>>(nB, m.bFaceCell.p, wallCell.p, iC.p, bC.p);
        keSetValuesCellK<<<blocks(n), kTpb, 0, c.stream>>>(n, wallCell.p, psi, Dm, S.p, psi);
        hipCheck(hipGetLastError(), "setValues");
    }
    const PcgPerf pf = pbicgstabSolve(c, m, LduView{n, nF, Dm, Up.p, Lo.p}, S.p, iC.p, bC.p, psi, ctl.solver);
    if (n) keBoundK<<<blocks(n), kTpb, 0, c.stream>>>(n, psi);
    const double* cells[1] = {psi};
    double* outs[1] = {psib};
    launchEvaluate(c, m, 1, 1, type, value, cells, outs);
    return pf;
}

KEpsilonResult KEpsilon::correct(const double* const* U, const double* const* Ub, const double* phi,
                                 const double* phiB, double* k, double* kb, double* eps, double* epsb, double* nut,
                                 double* nutb, double* gradScratch)
{
    const int n = m.nCells, nB = m.nBFaces;
    const KEpsilonCoeffs& co = ctl.co;
    KEpsilonResult res;
    res.nWallCells = nWallCells;
    // nut's boundary as it enters: the equations' patch diffusivities
    if (nB) launchCopy(c, nB, nutb, nutB0.p);
    // production from the current nut, grad(U) with every non-empty face at U's patch value, into the
    // caller's dead 9 n doubles when it lends them, else into the model's own (allocated on first use)
    double* gU = gradScratch;
    if (!gU)
    {
        gradU.resize(9 * (size_t)std::max(n, 1));
        gU = gradU.p;
    }
    {
        double* none[3] = {nullptr, nullptr, nullptr};   // only the packed gradient is read here
        launchGradGauss3(c, m, gType.p, Ub, g.Sf, g.w, g.V, U, none, gU);
    }
    if (n)
        keProductionK<<<blocks(n), kTpb, 0, c.stream>>>(n, m.ownerStart.p, m.losortStart.p, m.losort.p,
                                                        m.bndStart.p, m.bndFaces.p, isEmpty.p, gU, nut, phi,
                                                        phiB, g.V, gbn.p, G.p, divU.p);
    WallArgs wa{std::pow(co.Cmu, 0.25), std::pow(co.Cmu, 0.75), co.kappa, co.E, yPlusLamHost(co.kappa, co.E),
                ctl.nu};
    if (n)
        keWallK<<<blocks(n), kTpb, 0, c.stream>>>(n, m.bndStart.p, m.bndFaces.p, isWall.p, yWall.p, g.bDelta, U[0],
                                                  U[1], U[2], Ub[0], Ub[1], Ub[2], k, wa, G.p, eps, wallCell.p);
    hipCheck(hipGetLastError(), "kepsilon production");
    // epsilon first, then k with the new epsilon and the wall-overridden production
    const PcgPerf pe = solveScalar(true, phi, phiB, nut, eps, epsb, gbn.p, eps, k);
    res.epsInitial = pe.initialResidual;
    res.epsFinal = pe.finalResidual;
    res.epsIters = pe.nIterations;
    const PcgPerf pk = solveScalar(false, phi, phiB, nut, k, kb, G.p, eps, k);
    res.kInitial = pk.initialResidual;
    res.kFinal = pk.finalResidual;
    res.kIters = pk.nIterations;
    if (n) keNutK<<<blocks(n), kTpb, 0, c.stream>>>(n, co.Cmu, k, eps, nut);
    if (nB) keNutWallK<<<blocks(nB), kTpb, 0, c.stream>>>(nB, m.bFaceCell.p, isWall.p, yWall.p, k, wa, nutb);
    hipCheck(hipGetLastError(), "correctNut");
    hipCheck(hipStreamSynchronize(c.stream), "kepsilon correct");
    return res;
}

} // namespace mfv
