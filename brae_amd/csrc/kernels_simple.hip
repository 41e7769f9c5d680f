// The elementwise steps that compose one SIMPLE iteration on the device around the assembly, predictor and
// solver kernels (src/applications/solvers/simpleFoam/simpleFoam_cpp.cu:15-237 and the parts it calls:
// UEqn_cpp.cu:156-264, linearViscousStress_cpp.cu:105-184, pEqn_cpp.cu:31-300, fvc.cu:781-865). Each
// kernel restates its host loop's arithmetic term for term (the library is built without FMA contraction),
// so the composed step is the host step's to the bit when the solvers' reductions run sequentially.
//
// Boundary conditions per boundary face: kBcZeroGradient, kBcFixedValue (noSlip is a fixed value of 0),
// kBcEmpty (the face drops out of every sum).
#include "reduce.hpp"
#include "simple.hpp"

namespace mfv {

namespace {

constexpr int kTpb = 256;
inline int blocks(int n) { return (n + kTpb - 1) / kTpb; }

__global__ __launch_bounds__(kTpb) void fillK(int n, double v, double* __restrict__ x)
{
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i < n) x[i] = v;
}

__global__ __launch_bounds__(kTpb) void negateK(int n, const double* __restrict__ a, double* __restrict__ out)
{
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i < n) out[i] = -a[i];
}

// a += s*b elementwise (ldu_matrix.cuh:42-55 addEqual's loops)
__global__ __launch_bounds__(kTpb) void axpyK(int n, double s, const double* __restrict__ b, double* __restrict__ a)
{
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i < n) a[i] += s * b[i];
}

// the momentum matrix's boundary coefficients per face and component: fvm::div's (fvm.cuh:545-552;
// fixedValue: valueInternalCoeffs 0, valueBoundaryCoeffs value; zeroGradient: 1, 0) plus -1 times
// fvm::laplacian(nuEff)'s (fvm.cuh:43-103; fixedValue: gradientInternalCoeffs -deltaCoeffs,
// gradientBoundaryCoeffs deltaCoeffs*value), empty faces zero
__global__ __launch_bounds__(kTpb) void momBoundaryK(int nB, const int* __restrict__ type,
                                                     const double* __restrict__ value, const double* __restrict__ phib,
                                                     const double* __restrict__ nub, const double* __restrict__ bMagSf,
                                                     const double* __restrict__ bDelta, double* __restrict__ icX,
                                                     double* __restrict__ icY, double* __restrict__ icZ,
                                                     double* __restrict__ bcX, double* __restrict__ bcY,
                                                     double* __restrict__ bcZ)
{
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= nB) return;
    double* ic[3] = {icX, icY, icZ};
    double* bc[3] = {bcX, bcY, bcZ};
    const int t = type[i];
    if (t == kBcEmpty)
    {
        for (int k = 0; k < 3; ++k)
        {
            ic[k][i] = 0.0;
            bc[k][i] = 0.0;
        }
        return;
    }
    const double ph = phib[i];
    const double pGamma = nub[i] * bMagSf[i];
    for (int k = 0; k < 3; ++k)
    {
        double vIC = 1.0, vBC = 0.0, gIC = 0.0, gBC = 0.0;
        if (t == kBcFixedValue)
        {
            const double v = value[3 * (size_t)i + k];
            vIC = 0.0;
            vBC = v;
            gIC = -bDelta[i];
            gBC = bDelta[i] * v;
        }
        const double dIC = ph * vIC;
        const double dBC = (-ph) * vBC;
        const double lIC = pGamma * gIC;
        const double lBC = (-pGamma) * gBC;
        ic[k][i] = dIC + -1.0 * lIC;
        bc[k][i] = dBC + -1.0 * lBC;
    }
}

// patch values after evaluate() for up to three components (value stride 3, or 1 for a scalar field)
__global__ __launch_bounds__(kTpb) void evaluateK(int nB, int nComp, int stride, const label* __restrict__ bfc,
                                                  const int* __restrict__ type, const double* __restrict__ value,
                                                  const double* __restrict__ c0, const double* __restrict__ c1,
                                                  const double* __restrict__ c2, double* __restrict__ b0,
                                                  double* __restrict__ b1, double* __restrict__ b2)
{
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= nB) return;
    const double* cell[3] = {c0, c1, c2};
    double* out[3] = {b0, b1, b2};
    const int t = type[i];
    for (int k = 0; k < nComp; ++k)
    {
        double v = 0.0;
        if (t == kBcFixedValue)
            v = value[(size_t)stride * i + k];
        else if (t == kBcZeroGradient)
            v = cell[k][bfc[i]];
        out[k][i] = v;
    }
}

// three component gradients (xyz interleaved, 3n each) into grad(U) component-major: (i*3 + k) is dU_k/dx_i
__global__ __launch_bounds__(kTpb) void packGradK(int n, const double* __restrict__ g0, const double* __restrict__ g1,
                                                  const double* __restrict__ g2, double* __restrict__ g9)
{
    const int c = blockIdx.x * kTpb + threadIdx.x;
    if (c >= n) return;
    const double* g[3] = {g0, g1, g2};
    for (int k = 0; k < 3; ++k)
        for (int i = 0; i < 3; ++i) g9[(size_t)(i * 3 + k) * n + c] = g[k][3 * (size_t)c + i];
}

// the boundary-face stress nuEff_b*dev2(T(gb)) with gb = fvc::gradUBoundary (fvc.cu:781-865):
// gb = gc + n (sn - n.gc), n = (1/magSf) Sf, sn = deltaCoeffs*(Ub - Uc) on a fixedValue face, 0 on a
// zeroGradient one; dev2 as the cell stress (cf_types.cuh:174-179); empty faces zero
__global__ __launch_bounds__(kTpb) void stressBoundaryK(int n, int nF, int nB, const label* __restrict__ bfc,
                                                        const int* __restrict__ type, const double* __restrict__ Sf,
                                                        const double* __restrict__ magSf,
                                                        const double* __restrict__ bDelta,
                                                        const double* __restrict__ Ux, const double* __restrict__ Uy,
                                                        const double* __restrict__ Uz, const double* __restrict__ Ubx,
                                                        const double* __restrict__ Uby,
                                                        const double* __restrict__ Ubz,
                                                        const double* __restrict__ gradU,
                                                        const double* __restrict__ nub, double* __restrict__ sigB)
{
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= nB) return;
    const size_t sB = (size_t)nB;
    const int t = type[i];
    if (t == kBcEmpty)
    {
        for (int q = 0; q < 9; ++q) sigB[q * sB + i] = 0.0;
        return;
    }
    const int c = bfc[i];
    const size_t gf = (size_t)nF + i;
    const double r = 1.0 / magSf[gf];
    const double nv[3] = {r * Sf[3 * gf], r * Sf[3 * gf + 1], r * Sf[3 * gf + 2]};
    double gc[9];
    for (int q = 0; q < 9; ++q) gc[q] = gradU[(size_t)q * n + c];
    const double* U[3] = {Ux, Uy, Uz};
    const double* Ub[3] = {Ubx, Uby, Ubz};
    double sn[3];
    for (int j = 0; j < 3; ++j) sn[j] = t == kBcFixedValue ? bDelta[i] * (Ub[j][i] - U[j][c]) : 0.0;
    double gb[9];
    for (int j = 0; j < 3; ++j)
    {
        const double ndg = nv[0] * gc[0 * 3 + j] + nv[1] * gc[1 * 3 + j] + nv[2] * gc[2 * 3 + j];
        for (int a = 0; a < 3; ++a) gb[a * 3 + j] = gc[a * 3 + j] + nv[a] * (sn[j] - ndg);
    }
    const double tr = gb[0] + gb[4] + gb[8];
    const double kk = (2.0 / 3.0) * tr;
    const double nu = nub[i];
    for (int a = 0; a < 3; ++a)
        for (int j = 0; j < 3; ++j)
        {
            const double tt = gb[j * 3 + a];
            sigB[(size_t)(a * 3 + j) * sB + i] = nu * (a == j ? tt - kk : tt);
        }
}

// source_k -= (-(div_k/V))*V with div_k the stress divergence's raw face sum (linearViscousStress_cpp.cu:163-184)
__global__ __launch_bounds__(kTpb) void stressSourceK(int n, const double* __restrict__ V,
                                                      const double* __restrict__ dX, const double* __restrict__ dY,
                                                      const double* __restrict__ dZ, double* __restrict__ sX,
                                                      double* __restrict__ sY, double* __restrict__ sZ)
{
    const int c = blockIdx.x * kTpb + threadIdx.x;
    if (c >= n) return;
    const double v = V[c];
    const double e0 = -(dX[c] / v), e1 = -(dY[c] / v), e2 = -(dZ[c] / v);
    sX[c] -= e0 * v;
    sY[c] -= e1 * v;
    sZ[c] -= e2 * v;
}

// the momentum predictor's right-hand side: s_k = source_k - gradP_k*V (UEqn_cpp.cu:267-284)
__global__ __launch_bounds__(kTpb) void gradPSourceK(int n, const double* __restrict__ gradP, const double* __restrict__ V,
                                                     const double* __restrict__ s0, const double* __restrict__ s1,
                                                     const double* __restrict__ s2, double* __restrict__ o0,
                                                     double* __restrict__ o1, double* __restrict__ o2)
{
    const int c = blockIdx.x * kTpb + threadIdx.x;
    if (c >= n) return;
    const double v = V[c];
    o0[c] = s0[c] - gradP[3 * (size_t)c] * v;
    o1[c] = s1[c] - gradP[3 * (size_t)c + 1] * v;
    o2[c] = s2[c] - gradP[3 * (size_t)c + 2] * v;
}

// phiB = phiHbyA_b - (iC*p[cell] - bC) (fv_matrix_ops.cuh:75-78 matrixFlux on an uncoupled patch)
__global__ __launch_bounds__(kTpb) void boundaryFluxCorrK(int nB, const label* __restrict__ bfc,
                                                          const double* __restrict__ phiHbyAB,
                                                          const double* __restrict__ iC, const double* __restrict__ bC,
                                                          const double* __restrict__ p, double* __restrict__ phiB)
{
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= nB) return;
    phiB[i] = phiHbyAB[i] - (iC[i] * p[bfc[i]] - bC[i]);
}

// U = HbyA - rAU*grad(p) per component (pEqn_cpp.cu:279-292)
__global__ __launch_bounds__(kTpb) void correctUK(int n, const double* __restrict__ rAU, const double* __restrict__ gradP,
                                                  const double* __restrict__ h0, const double* __restrict__ h1,
                                                  const double* __restrict__ h2, double* __restrict__ u0,
                                                  double* __restrict__ u1, double* __restrict__ u2)
{
    const int c = blockIdx.x * kTpb + threadIdx.x;
    if (c >= n) return;
    const double r = rAU[c];
    u0[c] = h0[c] - r * gradP[3 * (size_t)c];
    u1[c] = h1[c] - r * gradP[3 * (size_t)c + 1];
    u2[c] = h2[c] - r * gradP[3 * (size_t)c + 2];
}

// a = a - b (the linearUpwind correction off the source, UEqn_cpp.cu:204-208)
__global__ __launch_bounds__(kTpb) void subInPlaceK(int n, const double* __restrict__ b, double* __restrict__ a)
{
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i < n) a[i] -= b[i];
}

// y = a*b elementwise (HbyA = rAU*H)
__global__ __launch_bounds__(kTpb) void mulK(int n, const double* __restrict__ a, const double* __restrict__ b,
                                             double* __restrict__ y)
{
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i < n) y[i] = a[i] * b[i];
}

} // namespace

void launchFill(Ctx& c, int n, double v, double* x)
{
    if (n <= 0) return;
    fillK<<<blocks(n), kTpb, 0, c.stream>>>(n, v, x);
    hipCheck(hipGetLastError(), "fill");
}

void launchNegate(Ctx& c, int n, const double* a, double* out)
{
    if (n <= 0) return;
    negateK<<<blocks(n), kTpb, 0, c.stream>>>(n, a, out);
    hipCheck(hipGetLastError(), "negate");
}

void launchAxpy(Ctx& c, int n, double s, const double* b, double* a)
{
    if (n <= 0) return;
    axpyK<<<blocks(n), kTpb, 0, c.stream>>>(n, s, b, a);
    hipCheck(hipGetLastError(), "axpy");
}

void launchMomBoundary(Ctx& c, int nB, const int* type, const double* value, const double* phib, const double* nub,
                       const double* bMagSf, const double* bDelta, double* const* iC, double* const* bC)
{
    if (nB <= 0) return;
    momBoundaryK<<<blocks(nB), kTpb, 0, c.stream>>>(nB, type, value, phib, nub, bMagSf, bDelta, iC[0], iC[1], iC[2],
                                                    bC[0], bC[1], bC[2]);
    hipCheck(hipGetLastError(), "momentum boundary coefficients");
}

void launchEvaluate(Ctx& c, const Mesh& m, int nComp, int stride, const int* type, const double* value,
                    const double* const* cell, double* const* out)
{
    if (m.nBFaces <= 0) return;
    evaluateK<<<blocks(m.nBFaces), kTpb, 0, c.stream>>>(m.nBFaces, nComp, stride, m.bFaceCell.p, type, value, cell[0],
                                                        nComp > 1 ? cell[1] : nullptr, nComp > 2 ? cell[2] : nullptr,
                                                        out[0], nComp > 1 ? out[1] : nullptr,
                                                        nComp > 2 ? out[2] : nullptr);
    hipCheck(hipGetLastError(), "evaluate boundary");
}

void launchPackGrad(Ctx& c, int n, const double* const* g, double* g9)
{
    if (n <= 0) return;
    packGradK<<<blocks(n), kTpb, 0, c.stream>>>(n, g[0], g[1], g[2], g9);
    hipCheck(hipGetLastError(), "pack gradient");
}

void launchStressBoundary(Ctx& c, const Mesh& m, const int* type, const double* Sf, const double* magSf,
                          const double* bDelta, const double* const* U, const double* const* Ub, const double* gradU,
                          const double* nub, double* sigB)
{
    if (m.nBFaces <= 0) return;
    stressBoundaryK<<<blocks(m.nBFaces), kTpb, 0, c.stream>>>(m.nCells, m.nFaces, m.nBFaces, m.bFaceCell.p, type, Sf,
                                                              magSf, bDelta, U[0], U[1], U[2], Ub[0], Ub[1], Ub[2],
                                                              gradU, nub, sigB);
    hipCheck(hipGetLastError(), "boundary stress");
}

void launchStressSource(Ctx& c, int n, const double* V, const double* const* d, double* const* src)
{
    if (n <= 0) return;
    stressSourceK<<<blocks(n), kTpb, 0, c.stream>>>(n, V, d[0], d[1], d[2], src[0], src[1], src[2]);
    hipCheck(hipGetLastError(), "stress source");
}

void launchGradPSource(Ctx& c, int n, const double* gradP, const double* V, const double* const* src,
                       double* const* out)
{
    if (n <= 0) return;
    gradPSourceK<<<blocks(n), kTpb, 0, c.stream>>>(n, gradP, V, src[0], src[1], src[2], out[0], out[1], out[2]);
    hipCheck(hipGetLastError(), "pressure gradient source");
}

void launchBoundaryFluxCorr(Ctx& c, const Mesh& m, const double* phiHbyAB, const double* iC, const double* bC,
                            const double* p, double* phiB)
{
    if (m.nBFaces <= 0) return;
    boundaryFluxCorrK<<<blocks(m.nBFaces), kTpb, 0, c.stream>>>(m.nBFaces, m.bFaceCell.p, phiHbyAB, iC, bC, p, phiB);
    hipCheck(hipGetLastError(), "boundary flux correction");
}

void launchCorrectU(Ctx& c, int n, const double* rAU, const double* gradP, const double* const* HbyA, double* const* U)
{
    if (n <= 0) return;
    correctUK<<<blocks(n), kTpb, 0, c.stream>>>(n, rAU, gradP, HbyA[0], HbyA[1], HbyA[2], U[0], U[1], U[2]);
    hipCheck(hipGetLastError(), "velocity correction");
}

void launchSubInPlace(Ctx& c, int n, const double* b, double* a)
{
    if (n <= 0) return;
    subInPlaceK<<<blocks(n), kTpb, 0, c.stream>>>(n, b, a);
    hipCheck(hipGetLastError(), "subtract");
}

void launchMul(Ctx& c, int n, const double* a, const double* b, double* y)
{
    if (n <= 0) return;
    mulK<<<blocks(n), kTpb, 0, c.stream>>>(n, a, b, y);
    hipCheck(hipGetLastError(), "multiply");
}

} // namespace mfv
