// LDU kernels on the mesh's internal-face addressing: the product A.x and the residual, the
// normFactor of OpenFOAM's residual norm, and the finite-volume assembly of fvm::laplacian and
// fvm::div (upwind) with the boundary fold, setReference and the matrix face flux.
//
// Every per-cell sum walks its faces in ASCENDING FACE INDEX, the order the reference's host loops
// accumulate in (ldu_spmv.cu:15-22, fvm.cuh:144-151). In OpenFOAM's upper-triangular face order a
// cell's neighbour-side faces (owner < cell) all precede the faces it owns, so that order is the
// cell's losort run followed by its owner run: a gather with no atomics whose FP64 result is the
// host loop's to the bit.
#include "reduce.hpp"

namespace mfv {

namespace {

constexpr int kTpb = 256;

inline int blocks(int n) { return (n + kTpb - 1) / kTpb; }

// y = A x (ldu_spmv.cu:5-24, device_spmv.cu:16-58)
__global__ __launch_bounds__(kTpb) void amulK(int n, const label* __restrict__ ownerStart,
                                              const label* __restrict__ losortStart,
                                              const label* __restrict__ losort,
                                              const label* __restrict__ losortOwner,
                                              const label* __restrict__ neighbour, const double* __restrict__ diag,
                                              const double* __restrict__ upper, const double* __restrict__ lower,
                                              const double* __restrict__ x, double* __restrict__ y)
{
    const int c = xcdBlock() * kTpb + threadIdx.x;
    if (c >= n) return;
    double s = diag[c] * x[c];
    const int l0 = losortStart[c], l1 = losortStart[c + 1];
    for (int k = l0; k < l1; ++k) s += lower[losort[k]] * x[losortOwner[k]];
    const int o0 = ownerStart[c], o1 = ownerStart[c + 1];
    for (int f = o0; f < o1; ++f) s += upper[f] * x[neighbour[f]];
    y[c] = s;
}

// r = b - A x, term by term (gamg_solver_cpp.cu:88-112)
__global__ __launch_bounds__(kTpb) void residualK(int n, const label* __restrict__ ownerStart,
                                                  const label* __restrict__ losortStart,
                                                  const label* __restrict__ losort,
                                                  const label* __restrict__ losortOwner,
                                                  const label* __restrict__ neighbour,
                                                  const double* __restrict__ diag, const double* __restrict__ upper,
                                                  const double* __restrict__ lower, const double* __restrict__ x,
                                                  const double* __restrict__ b, double* __restrict__ r)
{
    const int c = xcdBlock() * kTpb + threadIdx.x;
    if (c >= n) return;
    double s = b[c] - diag[c] * x[c];
    const int l0 = losortStart[c], l1 = losortStart[c + 1];
    for (int k = l0; k < l1; ++k) s -= lower[losort[k]] * x[losortOwner[k]];
    const int o0 = ownerStart[c], o1 = ownerStart[c + 1];
    for (int f = o0; f < o1; ++f) s -= upper[f] * x[neighbour[f]];
    r[c] = s;
}

__global__ __launch_bounds__(kReduceBlock) void sumK(int n, const double* __restrict__ a,
                                                     double* __restrict__ partials)
{
    __shared__ double lds[kReduceBlock / 64];
    double v = 0.0;
    MFV_ROWS(n, i) v += a[i];
    storePartial(blockSum(v, lds), partials);
}

__global__ void sumSeqK(int n, const double* __restrict__ a, double* __restrict__ out)
{
    if (threadIdx.x || blockIdx.x) return;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i];
    *out = s;
}

// normFactor summand |A.psi - sumA*xRef| + |b - sumA*xRef|, sumA = A.1 in the same face order
// (pcg.cu:68-97, device_pcg.cu:118-170); xRef = sum(psi)/n from the previous reduction.
struct NormPost
{
    double* out;
    __device__ void operator()(double s) const { *out = s + 1e-20; }
};

__device__ __forceinline__ double normSummand(int c, const label* __restrict__ ownerStart,
                                              const label* __restrict__ losortStart, const label* __restrict__ losort,
                                              const label* __restrict__ losortOwner,
                                              const label* __restrict__ neighbour, const double* __restrict__ diag,
                                              const double* __restrict__ upper, const double* __restrict__ lower,
                                              const double* __restrict__ psi, const double* __restrict__ b,
                                              double xRef)
{
    double ax = diag[c] * psi[c];
    double sa = diag[c];
    const int l0 = losortStart[c], l1 = losortStart[c + 1];
    for (int k = l0; k < l1; ++k)
    {
        const double lo = lower[losort[k]];
        ax += lo * psi[losortOwner[k]];
        sa += lo;
    }
    const int o0 = ownerStart[c], o1 = ownerStart[c + 1];
    for (int f = o0; f < o1; ++f)
    {
        ax += upper[f] * psi[neighbour[f]];
        sa += upper[f];
    }
    const double t = sa * xRef;
    return fabs(ax - t) + fabs(b[c] - t);
}

__global__ __launch_bounds__(kReduceBlock) void normFactorK(
    int n, const label* __restrict__ ownerStart, const label* __restrict__ losortStart,
    const label* __restrict__ losort, const label* __restrict__ losortOwner, const label* __restrict__ neighbour,
    const double* __restrict__ diag, const double* __restrict__ upper, const double* __restrict__ lower,
    const double* __restrict__ psi, const double* __restrict__ b, const double* __restrict__ sumPsi,
    double* __restrict__ partials)
{
    __shared__ double lds[kReduceBlock / 64];
    const double xRef = *sumPsi / (double)n;
    double v = 0.0;
    MFV_ROWS(n, c)
    v += normSummand(c, ownerStart, losortStart, losort, losortOwner, neighbour, diag, upper, lower, psi, b, xRef);
    storePartial(blockSum(v, lds), partials);
}

__global__ void normFactorSeqK(int n, const label* __restrict__ ownerStart, const label* __restrict__ losortStart,
                               const label* __restrict__ losort, const label* __restrict__ losortOwner,
                               const label* __restrict__ neighbour, const double* __restrict__ diag,
                               const double* __restrict__ upper, const double* __restrict__ lower,
                               const double* __restrict__ psi, const double* __restrict__ b,
                               const double* __restrict__ sumPsi, double* __restrict__ out)
{
    if (threadIdx.x || blockIdx.x) return;
    const double xRef = *sumPsi / (double)n;
    double s = 0.0;
    for (int c = 0; c < n; ++c)
        s += normSummand(c, ownerStart, losortStart, losort, losortOwner, neighbour, diag, upper, lower, psi, b, xRef);
    *out = s + 1e-20;
}

// fvm::laplacian face coefficients (fvm.cuh:144-148, device_fvm.cu:20-40)
__global__ __launch_bounds__(kTpb) void lapFaceK(int nF, const double* __restrict__ dc,
                                                 const double* __restrict__ gammaf, const double* __restrict__ magSf,
                                                 double* __restrict__ upper, double* __restrict__ lower)
{
    const int f = xcdBlock() * kTpb + threadIdx.x;
    if (f >= nF) return;
    const double coeff = dc[f] * gammaf[f] * magSf[f];
    upper[f] = coeff;
    if (lower) lower[f] = coeff;
}

// diag = -(sum of coefficients) in face order (fvm.cuh:148-149, device_fvm.cu:235-273)
__global__ __launch_bounds__(kTpb) void lapDiagK(int n, const label* __restrict__ ownerStart,
                                                 const label* __restrict__ losortStart,
                                                 const label* __restrict__ losort, const double* __restrict__ upper,
                                                 double* __restrict__ diag)
{
    const int c = xcdBlock() * kTpb + threadIdx.x;
    if (c >= n) return;
    double d = 0.0;
    const int l0 = losortStart[c], l1 = losortStart[c + 1];
    for (int k = l0; k < l1; ++k) d -= upper[losort[k]];
    const int o0 = ownerStart[c], o1 = ownerStart[c + 1];
    for (int f = o0; f < o1; ++f) d -= upper[f];
    diag[c] = d;
}

// fvm.cuh:88-101 for fixedValue (gradient coeffs -dc, dc*value), zeroGradient (0, 0), empty (none)
__global__ __launch_bounds__(kTpb) void lapBCoeffK(int nB, const int* __restrict__ bcType,
                                                   const double* __restrict__ bcValue,
                                                   const double* __restrict__ gammab,
                                                   const double* __restrict__ bMagSf,
                                                   const double* __restrict__ bDelta, double* __restrict__ iC,
                                                   double* __restrict__ bC)
{
    const int i = xcdBlock() * kTpb + threadIdx.x;
    if (i >= nB) return;
    const int t = bcType[i];
    if (t == 2)
    {
        iC[i] = 0.0;
        bC[i] = 0.0;
        return;
    }
    const double pGamma = gammab[i] * bMagSf[i];
    const double gIC = t == 1 ? -bDelta[i] : 0.0;
    const double gBC = t == 1 ? bDelta[i] * bcValue[i] : 0.0;
    iC[i] = pGamma * gIC;
    bC[i] = (-pGamma) * gBC;
}

// fvm::div(phi, psi) upwind faces (fvm.cuh:500-507, device_fvm.cu:64-76)
__global__ __launch_bounds__(kTpb) void divFaceK(int nF, const double* __restrict__ phi, double* __restrict__ upper,
                                                 double* __restrict__ lower)
{
    const int f = xcdBlock() * kTpb + threadIdx.x;
    if (f >= nF) return;
    const double ph = phi[f];
    const double w = ph >= 0.0 ? 1.0 : 0.0;
    const double lo = -w * ph;
    lower[f] = lo;
    upper[f] = lo + ph;
}

// fvm::div(phi, psi) Gauss linear faces: the geometric weight in place of pos0(phi)
// (divFaceLinearKernel, device_fvm.cu:43-61)
__global__ __launch_bounds__(kTpb) void divFaceLinearK(int nF, const double* __restrict__ phi,
                                                       const double* __restrict__ w, double* __restrict__ upper,
                                                       double* __restrict__ lower)
{
    const int f = xcdBlock() * kTpb + threadIdx.x;
    if (f >= nF) return;
    const double ph = phi[f];
    const double lo = -w[f] * ph;
    lower[f] = lo;
    upper[f] = lo + ph;
}

// diag -= lower (owned faces), -= upper (neighbour faces), in face order
__global__ __launch_bounds__(kTpb) void divDiagK(int n, const label* __restrict__ ownerStart,
                                                 const label* __restrict__ losortStart,
                                                 const label* __restrict__ losort, const double* __restrict__ upper,
                                                 const double* __restrict__ lower, double* __restrict__ diag)
{
    const int c = xcdBlock() * kTpb + threadIdx.x;
    if (c >= n) return;
    double d = 0.0;
    const int l0 = losortStart[c], l1 = losortStart[c + 1];
    for (int k = l0; k < l1; ++k) d -= upper[losort[k]];
    const int o0 = ownerStart[c], o1 = ownerStart[c + 1];
    for (int f = o0; f < o1; ++f) d -= lower[f];
    diag[c] = d;
}

// fvm.cuh:545-552: fixedValue value coeffs (0, value), zeroGradient (1, 0)
__global__ __launch_bounds__(kTpb) void divBCoeffK(int nB, const int* __restrict__ bcType,
                                                   const double* __restrict__ bcValue,
                                                   const double* __restrict__ phib, double* __restrict__ iC,
                                                   double* __restrict__ bC)
{
    const int i = xcdBlock() * kTpb + threadIdx.x;
    if (i >= nB) return;
    const int t = bcType[i];
    if (t == 2)
    {
        iC[i] = 0.0;
        bC[i] = 0.0;
        return;
    }
    const double vIC = t == 1 ? 0.0 : 1.0;
    const double vBC = t == 1 ? bcValue[i] : 0.0;
    iC[i] = phib[i] * vIC;
    bC[i] = (-phib[i]) * vBC;
}

// fvMatrix fold: diagC = diag + sum iC, b = source + sum bC over the cell's boundary faces in
// ascending face order (gamg_solver_cpp.cu:769-797, device_fvm.cu:866-912)
__global__ __launch_bounds__(kTpb) void foldK(int n, const label* __restrict__ bndStart,
                                              const label* __restrict__ bndFaces, const double* __restrict__ diag,
                                              const double* __restrict__ source, const double* __restrict__ iC,
                                              const double* __restrict__ bC, double* __restrict__ diagC,
                                              double* __restrict__ b)
{
    const int c = xcdBlock() * kTpb + threadIdx.x;
    if (c >= n) return;
    double d = diag[c], s = source[c];
    const int k0 = bndStart[c], k1 = bndStart[c + 1];
    for (int k = k0; k < k1; ++k)
    {
        const int i = bndFaces[k];
        d += iC[i];
        s += bC[i];
    }
    diagC[c] = d;
    b[c] = s;
}

// fvMatrix::setReference as the device path applies it (device_simple.cuh:104-106)
__global__ void setRefK(double* diag, double* b, int refCell, double refValue)
{
    if (threadIdx.x || blockIdx.x) return;
    b[refCell] += diag[refCell] * refValue;
    diag[refCell] += diag[refCell];
}

// pEqn.flux(): upper*psi[nei] - lower*psi[own] (device_simple.cuh:51-52, pEqn_cpp.cuh:134-139)
__global__ __launch_bounds__(kTpb) void faceFluxK(int nF, const label* __restrict__ owner,
                                                  const label* __restrict__ neighbour,
                                                  const double* __restrict__ upper, const double* __restrict__ lower,
                                                  const double* __restrict__ psi, double* __restrict__ flux)
{
    const int f = xcdBlock() * kTpb + threadIdx.x;
    if (f >= nF) return;
    flux[f] = upper[f] * psi[neighbour[f]] - lower[f] * psi[owner[f]];
}

// linear interpolation w*vf[own] + (1 - w)*vf[nei] (fvc.cu:626-656); boundary faces take the owner
// cell's value (an extrapolated field: rAU's fvc::interpolate, pEqn_cpp.cu:217-222)
__global__ __launch_bounds__(kTpb) void interpK(int nF, int nB, const label* __restrict__ owner,
                                                const label* __restrict__ neighbour,
                                                const label* __restrict__ bFaceCell, const double* __restrict__ w,
                                                const double* __restrict__ vol, double* __restrict__ face,
                                                double* __restrict__ bface)
{
    const int f = xcdBlock() * kTpb + threadIdx.x;
    if (f < nF) face[f] = w[f] * vol[owner[f]] + (1.0 - w[f]) * vol[neighbour[f]];
    else if (f < nF + nB && bface) bface[f - nF] = vol[bFaceCell[f - nF]];
}

// source += sign * (div(phi) * V) with div = (sum_own phi - sum_nei phi + sum_boundary phi_b)/V in face
// order (fvc.cu:706-730; pEqn_cpp.cu:238-242)
__global__ __launch_bounds__(kTpb) void divSourceK(int n, const label* __restrict__ ownerStart,
                                                   const label* __restrict__ losortStart,
                                                   const label* __restrict__ losort,
                                                   const label* __restrict__ bndStart,
                                                   const label* __restrict__ bndFaces, const double* __restrict__ phi,
                                                   const double* __restrict__ phib, const double* __restrict__ V,
                                                   double sign, double* __restrict__ source)
{
    const int c = xcdBlock() * kTpb + threadIdx.x;
    if (c >= n) return;
    double d = 0.0;
    const int l0 = losortStart[c], l1 = losortStart[c + 1];
    for (int k = l0; k < l1; ++k) d -= phi[losort[k]];
    const int o0 = ownerStart[c], o1 = ownerStart[c + 1];
    for (int f = o0; f < o1; ++f) d += phi[f];
    if (phib)
        for (int k = bndStart[c]; k < bndStart[c + 1]; ++k) d += phib[bndFaces[k]];
    d /= V[c];
    source[c] += sign * (d * V[c]);
}

// phi = phiHbyA - sign*(upper*psi[nei] - lower*psi[own] + ffc) (pEqn_cpp.cuh:134-139 with the
// matrix's sign; ffc the stored non-orthogonal correction flux, fv_matrix_ops.cuh:46-48, or none)
__global__ __launch_bounds__(kTpb) void fluxCorrectK(int nF, const label* __restrict__ owner,
                                                     const label* __restrict__ neighbour,
                                                     const double* __restrict__ upper,
                                                     const double* __restrict__ lower, const double* __restrict__ psi,
                                                     const double* __restrict__ ffc,
                                                     const double* __restrict__ phiHbyA, double sign,
                                                     double* __restrict__ phi)
{
    const int f = xcdBlock() * kTpb + threadIdx.x;
    if (f >= nF) return;
    double flux = upper[f] * psi[neighbour[f]] - lower[f] * psi[owner[f]];
    if (ffc) flux += ffc[f];
    phi[f] = phiHbyA[f] - sign * flux;
}

// ---------------- non-orthogonal correction ----------------

// ffc = (gamma*magSf) * (corrVec . (w*grad[own] + (1 - w)*grad[nei])) per internal face (fvm.cuh:170-222)
__global__ __launch_bounds__(kTpb) void corrFluxK(int nF, const label* __restrict__ owner,
                                                  const label* __restrict__ neighbour, const double* __restrict__ grad,
                                                  const double* __restrict__ w, const double* __restrict__ cv,
                                                  const double* __restrict__ gammaf, const double* __restrict__ magSf,
                                                  double* __restrict__ ffc)
{
    const int f = xcdBlock() * kTpb + threadIdx.x;
    if (f >= nF) return;
    const int o = owner[f], nb = neighbour[f];
    const double wf = w[f], wn = 1.0 - wf;
    const double fx = wf * grad[3 * o] + wn * grad[3 * nb];
    const double fy = wf * grad[3 * o + 1] + wn * grad[3 * nb + 1];
    const double fz = wf * grad[3 * o + 2] + wn * grad[3 * nb + 2];
    const double corr = cv[3 * f] * fx + cv[3 * f + 1] * fy + cv[3 * f + 2] * fz;
    ffc[f] = (gammaf[f] * magSf[f]) * corr;
}

// `limited k corrected` (limitedSnGrad; fvm.cuh:170-211, lapCorrFaceLimitedKernel device_fvm.cu:395-437): the
// correction capped per face against the orthogonal snGrad, lim = min(k|orth| / ((1-k)|corr| + 1e-15), 1)
__global__ __launch_bounds__(kTpb) void corrFluxLimitedK(int nF, const label* __restrict__ owner,
                                                         const label* __restrict__ neighbour,
                                                         const double* __restrict__ grad, const double* __restrict__ w,
                                                         const double* __restrict__ cv,
                                                         const double* __restrict__ gammaf,
                                                         const double* __restrict__ magSf,
                                                         const double* __restrict__ nonOrthDc,
                                                         const double* __restrict__ vf, double k,
                                                         double* __restrict__ ffc)
{
    const int f = xcdBlock() * kTpb + threadIdx.x;
    if (f >= nF) return;
    const int o = owner[f], nb = neighbour[f];
    const double wf = w[f], wn = 1.0 - wf;
    const double fx = wf * grad[3 * o] + wn * grad[3 * nb];
    const double fy = wf * grad[3 * o + 1] + wn * grad[3 * nb + 1];
    const double fz = wf * grad[3 * o + 2] + wn * grad[3 * nb + 2];
    const double corr = cv[3 * f] * fx + cv[3 * f + 1] * fy + cv[3 * f + 2] * fz;
    const double orth = nonOrthDc[f] * (vf[nb] - vf[o]);
    const double lim = fmin(k * fabs(orth) / ((1.0 - k) * fabs(corr) + 1e-15), 1.0);
    ffc[f] = (gammaf[f] * magSf[f]) * (lim * corr);
}

// source -= sum over the cell's faces of (+ffc owned, -ffc neighbour-side), ascending face index
// (fvm.cuh:293-299, pEqn_cpp.cu:228)
__global__ __launch_bounds__(kTpb) void corrSourceK(int n, const label* __restrict__ ownerStart,
                                                    const label* __restrict__ losortStart,
                                                    const label* __restrict__ losort, const double* __restrict__ ffc,
                                                    double* __restrict__ source)
{
    const int c = xcdBlock() * kTpb + threadIdx.x;
    if (c >= n) return;
    double s = 0.0;
    for (int k = losortStart[c]; k < losortStart[c + 1]; ++k) s -= ffc[losort[k]];
    for (int f = ownerStart[c]; f < ownerStart[c + 1]; ++f) s += ffc[f];
    source[c] -= s;
}

} // namespace

void launchInterpolate(Ctx& c, const Mesh& m, const double* w, const double* vol, double* face, double* bface)
{
    const int tot = m.nFaces + (bface ? m.nBFaces : 0);
    if (!tot) return;
    interpK<<<blocks(tot), kTpb, 0, c.stream>>>(m.nFaces, m.nBFaces, m.owner.p, m.neighbour.p, m.bFaceCell.p, w, vol,
                                                face, bface);
    hipCheck(hipGetLastError(), "interpolate");
}

void launchDivSource(Ctx& c, const Mesh& m, const double* phi, const double* phib, const double* V, double sign,
                     double* source)
{
    if (!m.nCells) return;
    divSourceK<<<blocks(m.nCells), kTpb, 0, c.stream>>>(m.nCells, m.ownerStart.p, m.losortStart.p, m.losort.p,
                                                        m.bndStart.p, m.bndFaces.p, phi, phib, V, sign, source);
    hipCheck(hipGetLastError(), "div source");
}

void launchFluxCorrect(Ctx& c, const Mesh& m, const LduView& A, const double* psi, const double* ffc,
                       const double* phiHbyA, double sign, double* phi)
{
    if (!m.nFaces) return;
    fluxCorrectK<<<blocks(m.nFaces), kTpb, 0, c.stream>>>(m.nFaces, m.owner.p, m.neighbour.p, A.upper, A.lower, psi,
                                                          ffc, phiHbyA, sign, phi);
    hipCheck(hipGetLastError(), "flux correct");
}

void launchLaplacianCorr(Ctx& c, const Mesh& m, const double* grad, const double* w, const double* corrVec,
                         const double* gammaf, const double* magSf, double* ffc, double* source,
                         const double* nonOrthDc, const double* vf, double limitCoeff)
{
    const bool limited = limitCoeff > 0.0 && limitCoeff < 1.0;
    if (m.nFaces && limited)
        corrFluxLimitedK<<<blocks(m.nFaces), kTpb, 0, c.stream>>>(m.nFaces, m.owner.p, m.neighbour.p, grad, w, corrVec,
                                                                  gammaf, magSf, nonOrthDc, vf, limitCoeff, ffc);
    else if (m.nFaces)
        corrFluxK<<<blocks(m.nFaces), kTpb, 0, c.stream>>>(m.nFaces, m.owner.p, m.neighbour.p, grad, w, corrVec, gammaf,
                                                           magSf, ffc);
    if (source && m.nCells)
        corrSourceK<<<blocks(m.nCells), kTpb, 0, c.stream>>>(m.nCells, m.ownerStart.p, m.losortStart.p, m.losort.p,
                                                             ffc, source);
    hipCheck(hipGetLastError(), "laplacian correction");
}

void launchAmul(Ctx& c, const Mesh& m, const LduView& A, const double* x, double* y)
{
    if (m.nCells == 0) return;
    amulK<<<blocks(m.nCells), kTpb, 0, c.stream>>>(m.nCells, m.ownerStart.p, m.losortStart.p, m.losort.p,
                                                   m.losortOwner.p, m.neighbour.p, A.diag, A.upper, A.lower, x, y);
    hipCheck(hipGetLastError(), "amul");
}

void launchResidual(Ctx& c, const Mesh& m, const LduView& A, const double* x, const double* b, double* r)
{
    if (m.nCells == 0) return;
    residualK<<<blocks(m.nCells), kTpb, 0, c.stream>>>(m.nCells, m.ownerStart.p, m.losortStart.p, m.losort.p,
                                                       m.losortOwner.p, m.neighbour.p, A.diag, A.upper, A.lower, x,
                                                       b, r);
    hipCheck(hipGetLastError(), "residual");
}

void launchNormFactor(Ctx& c, const Mesh& m, const LduView& A, const double* psi, const double* b, double* out)
{
    const int n = m.nCells;
    double* sumPsi = c.scalars.p + 63;
    if (c.sequentialReduce)
    {
        sumSeqK<<<1, 64, 0, c.stream>>>(n, psi, sumPsi);
        normFactorSeqK<<<1, 64, 0, c.stream>>>(n, m.ownerStart.p, m.losortStart.p, m.losort.p, m.losortOwner.p,
                                               m.neighbour.p, A.diag, A.upper, A.lower, psi, b, sumPsi, out);
    }
    else
    {
        const int g = reduceGrid(n);
        double* pa = c.partials.p + 3 * (size_t)kMaxReduceBlocks;
        sumK<<<g, kReduceBlock, 0, c.stream>>>(n, psi, pa);
        finishK<<<1, kFinishBlock, 0, c.stream>>>(g, pa, StoreTo{sumPsi}, nullptr);
        normFactorK<<<g, kReduceBlock, 0, c.stream>>>(n, m.ownerStart.p, m.losortStart.p, m.losort.p,
                                                      m.losortOwner.p, m.neighbour.p, A.diag, A.upper, A.lower, psi,
                                                      b, sumPsi, pa);
        finishK<<<1, kFinishBlock, 0, c.stream>>>(g, pa, NormPost{out}, nullptr);
    }
    hipCheck(hipGetLastError(), "normFactor");
}

void launchLaplacian(Ctx& c, const Mesh& m, const double* dc, const double* gammaf, const double* magSf,
                     double* diag, double* upper, double* lower)
{
    if (m.nFaces)
        lapFaceK<<<blocks(m.nFaces), kTpb, 0, c.stream>>>(m.nFaces, dc, gammaf, magSf, upper,
                                                          lower == upper ? nullptr : lower);
    if (m.nCells)
        lapDiagK<<<blocks(m.nCells), kTpb, 0, c.stream>>>(m.nCells, m.ownerStart.p, m.losortStart.p, m.losort.p,
                                                          upper, diag);
    hipCheck(hipGetLastError(), "laplacian");
}

void launchLaplacianBCoeffs(Ctx& c, int nB, const int* bcType, const double* bcValue, const double* gammab,
                            const double* bMagSf, const double* bDelta, double* iC, double* bC)
{
    if (!nB) return;
    lapBCoeffK<<<blocks(nB), kTpb, 0, c.stream>>>(nB, bcType, bcValue, gammab, bMagSf, bDelta, iC, bC);
    hipCheck(hipGetLastError(), "laplacian boundary coefficients");
}

void launchDivUpwind(Ctx& c, const Mesh& m, const double* phi, double* diag, double* upper, double* lower)
{
    if (m.nFaces) divFaceK<<<blocks(m.nFaces), kTpb, 0, c.stream>>>(m.nFaces, phi, upper, lower);
    if (m.nCells)
        divDiagK<<<blocks(m.nCells), kTpb, 0, c.stream>>>(m.nCells, m.ownerStart.p, m.losortStart.p, m.losort.p,
                                                          upper, lower, diag);
    hipCheck(hipGetLastError(), "div upwind");
}

void launchDivCentral(Ctx& c, const Mesh& m, const double* phi, const double* weights, double* diag, double* upper,
                      double* lower)
{
    if (m.nFaces) divFaceLinearK<<<blocks(m.nFaces), kTpb, 0, c.stream>>>(m.nFaces, phi, weights, upper, lower);
    if (m.nCells)
        divDiagK<<<blocks(m.nCells), kTpb, 0, c.stream>>>(m.nCells, m.ownerStart.p, m.losortStart.p, m.losort.p,
                                                          upper, lower, diag);
    hipCheck(hipGetLastError(), "div linear");
}

void launchDivBCoeffs(Ctx& c, int nB, const int* bcType, const double* bcValue, const double* phib, double* iC,
                      double* bC)
{
    if (!nB) return;
    divBCoeffK<<<blocks(nB), kTpb, 0, c.stream>>>(nB, bcType, bcValue, phib, iC, bC);
    hipCheck(hipGetLastError(), "div boundary coefficients");
}

void launchFold(Ctx& c, const Mesh& m, const double* diag, const double* source, const double* iC, const double* bC,
                double* diagC, double* b)
{
    if (!m.nCells) return;
    foldK<<<blocks(m.nCells), kTpb, 0, c.stream>>>(m.nCells, m.bndStart.p, m.bndFaces.p, diag, source, iC, bC, diagC,
                                                   b);
    hipCheck(hipGetLastError(), "fold");
}

void launchSetReference(Ctx& c, double* diag, double* b, int refCell, double refValue)
{
    setRefK<<<1, 64, 0, c.stream>>>(diag, b, refCell, refValue);
    hipCheck(hipGetLastError(), "setReference");
}

void launchFaceFlux(Ctx& c, const Mesh& m, const LduView& A, const double* psi, double* flux)
{
    if (!m.nFaces) return;
    faceFluxK<<<blocks(m.nFaces), kTpb, 0, c.stream>>>(m.nFaces, m.owner.p, m.neighbour.p, A.upper, A.lower, psi,
                                                       flux);
    hipCheck(hipGetLastError(), "face flux");
}

} // namespace mfv
