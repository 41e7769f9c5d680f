// The SIMPLE pressure predictor's momentum glue on the device: A()'s folded diagonal and rAU, H() per
// component (and HbyA = rAU*H), constrainHbyA, phiHbyA = flux(HbyA) with its boundary faces, and
// adjustPhi. Reference: pressurePredictor (src/applications/solvers/simpleFoam/pEqn.cu:94-335) and its
// host twin (pEqn_cpp.cu:40-133, fv_matrix_ops.cuh:211-340, fvc.cu:545-580).
//
// Every per-cell sum is a gather in the host loop's order, term by term: a cell's boundary faces
// ascending (the patch order the host walks), and its internal faces ascending -- the losort run, then
// the owned run (OpenFOAM's upper-triangular order puts every face a cell neighbours before the faces it
// owns). No atomics, so the FP64 results are the host loops' to the bit.
#include "reduce.hpp"

namespace mfv {

namespace {

constexpr int kTpb = 256;
static_assert(kTpb == kReduceBlock, "slabRow numbers threads of kReduceBlock-wide workgroups");
inline int blocks(int n) { return (n + kTpb - 1) / kTpb; }

constexpr int kLinUpSplit = 2;   // linearUpwindCorrK's column parts

// D = diag + cmptAv(internalCoeffs) face by face (fv_matrix_ops.cuh:211-226; pEqn.cu:117-138, the
// cmptAv and foldBoundaryDiag kernels), cmptAv = (x + y + z)/3
__global__ __launch_bounds__(kTpb) void foldBoundaryDiagK(int n, const label* __restrict__ bndStart,
                                                          const label* __restrict__ bndFaces,
                                                          const double* __restrict__ icX,
                                                          const double* __restrict__ icY,
                                                          const double* __restrict__ icZ,
                                                          const double* __restrict__ diag, double* __restrict__ D)
{
    const int c = xcdBlock() * kTpb + threadIdx.x;
    if (c >= n) return;
    double d = diag[c];
    for (int k = bndStart[c]; k < bndStart[c + 1]; ++k)
    {
        const int i = bndFaces[k];
        d += (icX[i] + icY[i] + icZ[i]) / 3.0;
    }
    D[c] = d;
}

// rAU = 1/A, A = D/V (device_simple.cu reciprocalVKernel; fv_matrix_ops.cuh:224-225)
__global__ __launch_bounds__(kTpb) void reciprocalVK(int n, const double* __restrict__ D, const double* __restrict__ V,
                                                     double* __restrict__ rAU)
{
    const int c = xcdBlock() * kTpb + threadIdx.x;
    if (c >= n) return;
    const double A = D[c] / V[c];
    rAU[c] = 1.0 / A;
}

// cmptAv(ic) - ic_k per boundary face (pEqn.cu:146-149; fv_matrix_ops.cuh:290-300)
__global__ __launch_bounds__(kTpb) void boundaryDiagTermK(int nB, const double* __restrict__ icX,
                                                          const double* __restrict__ icY,
                                                          const double* __restrict__ icZ, int k,
                                                          double* __restrict__ out)
{
    const int i = xcdBlock() * kTpb + threadIdx.x;
    if (i >= nB) return;
    const double x = icX[i], y = icY[i], z = icZ[i];
    const double av = (x + y + z) / 3.0;
    out[i] = av - (k == 0 ? x : (k == 1 ? y : z));
}

// H() of one component (fv_matrix_ops.cuh:277-340; device_simple.cu matrixHKernel), optionally times rAU
// (HbyA, pEqn_cpp.cu:52-53): from zero, bdDiag*psi per boundary face, minus lower*psi[owner] and
// upper*psi[neighbour] per internal face, plus the source, plus boundaryCoeffs per boundary face, over V
template <bool Slab>
__global__ __launch_bounds__(kTpb) void matrixHK(int n, SlabPart part, const label* __restrict__ ownerStart,
                                                 const label* __restrict__ losortStart,
                                                 const label* __restrict__ losort,
                                                 const label* __restrict__ losortOwner,
                                                 const label* __restrict__ neighbour,
                                                 const label* __restrict__ bndStart,
                                                 const label* __restrict__ bndFaces,
                                                 const double* __restrict__ upper, const double* __restrict__ lower,
                                                 const double* __restrict__ psi, const double* __restrict__ source,
                                                 const double* __restrict__ bdDiag, const double* __restrict__ bdSrc,
                                                 const double* __restrict__ V, const double* __restrict__ rAU,
                                                 double* __restrict__ out)
{
    const int c = gatherCell<Slab>(n, part);
    if (c < 0) return;
    const int b0 = bndStart[c], b1 = bndStart[c + 1];
    const double pc = psi[c];
    double h = 0.0;
    for (int k = b0; k < b1; ++k) h += bdDiag[bndFaces[k]] * pc;
    for (int k = losortStart[c]; k < losortStart[c + 1]; ++k) h = h - lower[losort[k]] * psi[losortOwner[k]];
    for (int f = ownerStart[c]; f < ownerStart[c + 1]; ++f) h = h - upper[f] * psi[neighbour[f]];
    h += source[c];
    for (int k = b0; k < b1; ++k) h += bdSrc[bndFaces[k]];
    h = h / V[c];
    out[c] = rAU ? rAU[c] * h : h;
}

// The three components of a vector field at once: one pass over the matrix and the addressing
struct Vec3Args
{
    const double* psi[3];
    const double* source[3];
    const double* bdDiag[3];
    const double* bdSrc[3];
    double* out[3];
    int valid[3];
};

// H() (or HbyA) of the three components in one pass: each component's terms in matrixHK's order, so the
// values are three matrixHK launches' to the bit; a component that is not valid is written as zeros.
// The first kBatch faces of the cell's losort run and of its owned run are loaded together, each level of
// the gather (face index, coefficient and neighbour, psi) at once for both runs: clamped indices keep
// every load in bounds and unconditional, and a face beyond its run is dropped from the sum, not added as
// zero. Faces past the first kBatch of a run follow in a loop.
constexpr int kBatch = 3;

template <bool Slab>
__global__ __launch_bounds__(kTpb) void matrixH3K(int n, SlabPart part, int nF, const label* __restrict__ ownerStart,
                                                  const label* __restrict__ losortStart,
                                                  const label* __restrict__ losort,
                                                  const label* __restrict__ losortOwner,
                                                  const label* __restrict__ neighbour,
                                                  const label* __restrict__ bndStart,
                                                  const label* __restrict__ bndFaces,
                                                  const double* __restrict__ upper, const double* __restrict__ lower,
                                                  Vec3Args a, const double* __restrict__ V,
                                                  const double* __restrict__ rAU)
{
    const int c = gatherCell<Slab>(n, part);
    if (c < 0) return;
    const int b0 = bndStart[c], b1 = bndStart[c + 1];
    const int l0 = losortStart[c], l1 = losortStart[c + 1];
    const int f0 = ownerStart[c], f1 = ownerStart[c + 1];
    double h[3], pc[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        h[k] = 0.0;
        pc[k] = a.psi[k][c];
    }
    double lv[kBatch] = {}, uv[kBatch] = {}, lp[kBatch][3] = {}, up[kBatch][3] = {};
    if (nF > 0)   // uniform: a mesh with internal faces
    {
        const int last = nF - 1;
        int lf[kBatch], lo[kBatch], nb[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
        {
            const int j = min(l0 + u, last), f = min(f0 + u, last);
            lf[u] = losort[j];
            lo[u] = losortOwner[j];
            nb[u] = neighbour[f];
            uv[u] = upper[f];
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
        {
            lv[u] = lower[lf[u]];
#pragma unroll
            for (int k = 0; k < 3; ++k)
            {
                lp[u][k] = a.psi[k][lo[u]];
                up[u][k] = a.psi[k][nb[u]];
            }
        }
    }
    for (int j = b0; j < b1; ++j)
    {
        const int i = bndFaces[j];
#pragma unroll
        for (int k = 0; k < 3; ++k) h[k] += a.bdDiag[k][i] * pc[k];
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (l0 + u < l1) h[k] = h[k] - lv[u] * lp[u][k];
    for (int j = l0 + kBatch; j < l1; ++j)
    {
        const double l = lower[losort[j]];
        const int o = losortOwner[j];
#pragma unroll
        for (int k = 0; k < 3; ++k) h[k] = h[k] - l * a.psi[k][o];
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (f0 + u < f1) h[k] = h[k] - uv[u] * up[u][k];
    for (int f = f0 + kBatch; f < f1; ++f)
    {
        const double u = upper[f];
        const int nbf = neighbour[f];
#pragma unroll
        for (int k = 0; k < 3; ++k) h[k] = h[k] - u * a.psi[k][nbf];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) h[k] += a.source[k][c];
    for (int j = b0; j < b1; ++j)
    {
        const int i = bndFaces[j];
#pragma unroll
        for (int k = 0; k < 3; ++k) h[k] += a.bdSrc[k][i];
    }
    const double v = V[c];
    const double r = rAU ? rAU[c] : 1.0;
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        const double hk = h[k] / v;
        a.out[k][c] = !a.valid[k] ? 0.0 : (rAU ? r * hk : hk);
    }
}

// fvMatrix::relax for uncoupled patches (fv_matrix_ops.cuh:158-207; device_simple.cu relaxKernel) per cell,
// in the host loops' order: sumOff = |lower| over the losort run then |upper| over the owned run (face
// order), plus the coupled-interface magnitude; d = diag + cmptMagMax(ic) face by face; d = max(|d|,
// sumOff)/alpha; d -= cmptMin(ic) face by face; delta = d - diag; and, for nComp components, source_k +=
// delta*psi_k in the same pass. icMag/icMin NULL: |icBnd| and icBnd (equal components).
struct RelaxArgs
{
    const double* icBnd;
    const double* icMag;
    const double* icMin;
    const double* cycSumOff;
    double alpha;
    int nComp;
    const double* psi[3];
    double* source[3];
};

template <bool Slab>
__global__ __launch_bounds__(kTpb) void relaxK(int n, SlabPart part, int nF, const label* __restrict__ ownerStart,
                                               const label* __restrict__ losortStart,
                                               const label* __restrict__ losort, const label* __restrict__ bndStart,
                                               const label* __restrict__ bndFaces, const double* __restrict__ upper,
                                               const double* __restrict__ lower, const double* __restrict__ diag,
                                               RelaxArgs a, double* __restrict__ relaxed, double* __restrict__ delta)
{
    const int c = gatherCell<Slab>(n, part);
    if (c < 0) return;
    const int b0 = bndStart[c], b1 = bndStart[c + 1];
    const int l0 = losortStart[c], l1 = losortStart[c + 1];
    const int f0 = ownerStart[c], f1 = ownerStart[c + 1];
    double lv[kBatch] = {}, uv[kBatch] = {};
    if (nF > 0)   // uniform
    {
        const int last = nF - 1;
        int lf[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
        {
            lf[u] = losort[min(l0 + u, last)];
            uv[u] = upper[min(f0 + u, last)];
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) lv[u] = lower[lf[u]];
    }
    const double d0 = diag[c];
    double sumOff = 0.0;
#pragma unroll
    for (int u = 0; u < kBatch; ++u)
        if (l0 + u < l1) sumOff += fabs(lv[u]);
    for (int j = l0 + kBatch; j < l1; ++j) sumOff += fabs(lower[losort[j]]);
#pragma unroll
    for (int u = 0; u < kBatch; ++u)
        if (f0 + u < f1) sumOff += fabs(uv[u]);
    for (int f = f0 + kBatch; f < f1; ++f) sumOff += fabs(upper[f]);
    if (a.cycSumOff) sumOff += a.cycSumOff[c];
    double d = d0;
    for (int j = b0; j < b1; ++j)
    {
        const int i = bndFaces[j];
        d += a.icMag ? a.icMag[i] : fabs(a.icBnd[i]);
    }
    d = fmax(fabs(d), sumOff) / a.alpha;
    for (int j = b0; j < b1; ++j)
    {
        const int i = bndFaces[j];
        d -= a.icMin ? a.icMin[i] : a.icBnd[i];
    }
    const double dd = d - d0;
    if (relaxed) relaxed[c] = d;
    if (delta) delta[c] = dd;
    for (int k = 0; k < a.nComp; ++k) a.source[k][c] += dd * a.psi[k][c];
}

// linearUpwind's deferred correction of div(phi, U_k) for NC components (fvm.cuh:395-433;
// device_deferred_correction.cu linearUpwindCorrKernel): per cell, over its internal faces in face order
// (the losort run, then the owned run), fc = phi * ((Cf - C_up) . grad_k(up)) with up the owner where
// phi > 0, subtracted on a neighbour's face and added on an owned one. The dot product is x + y + z,
// uncontracted, as the host source writes it. The first kBatch faces of each run are loaded together
// (clamped indices, dropped from the sum past the run's end), the rest in a loop.
struct LinUpArgs
{
    const double* grad[3];   // xyz per cell
    double* corr[3];
};

// one face's fc for NC components: the upwind cell's centre and gradient are selected between this
// cell's (held in registers) and the other cell's, both loaded before phi is known
template <int NC>
struct LinUpCell
{
    double x, y, z, g[NC][3];
};

template <int NC>
__device__ __forceinline__ void linUpLoad(int cell, const double* __restrict__ C, const LinUpArgs& a,
                                          LinUpCell<NC>& o)
{
    o.x = C[3 * (size_t)cell];
    o.y = C[3 * (size_t)cell + 1];
    o.z = C[3 * (size_t)cell + 2];
#pragma unroll
    for (int k = 0; k < NC; ++k)
#pragma unroll
        for (int i = 0; i < 3; ++i) o.g[k][i] = a.grad[k][3 * (size_t)cell + i];
}

template <int NC>
__device__ __forceinline__ void linUpFace(double pf, bool otherUp, const LinUpCell<NC>& self,
                                          const LinUpCell<NC>& other, double fx, double fy, double fz,
                                          double (&fc)[NC])
{
    const double dx = fx - (otherUp ? other.x : self.x);
    const double dy = fy - (otherUp ? other.y : self.y);
    const double dz = fz - (otherUp ? other.z : self.z);
#pragma unroll
    for (int k = 0; k < NC; ++k)
    {
        const double gx = otherUp ? other.g[k][0] : self.g[k][0];
        const double gy = otherUp ? other.g[k][1] : self.g[k][1];
        const double gz = otherUp ? other.g[k][2] : self.g[k][2];
        fc[k] = pf * (dx * gx + dy * gy + dz * gz);
    }
}

template <bool Slab, int NC>
__global__ __launch_bounds__(kTpb) void linearUpwindCorrK(int n, SlabPart part, int nF,
                                                          const label* __restrict__ ownerStart,
                                                          const label* __restrict__ losortStart,
                                                          const label* __restrict__ losort,
                                                          const label* __restrict__ losortOwner,
                                                          const label* __restrict__ neighbour,
                                                          const double* __restrict__ phi,
                                                          const double* __restrict__ Cf,
                                                          const double* __restrict__ C, LinUpArgs a)
{
    const int c = gatherCell<Slab, kLinUpSplit>(n, part);
    if (c < 0) return;
    const int l0 = losortStart[c], l1 = losortStart[c + 1];
    const int f0 = ownerStart[c], f1 = ownerStart[c + 1];
    LinUpCell<NC> self;
    linUpLoad<NC>(c, C, a, self);
    double s[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) s[k] = 0.0;
    // the losort run (this cell the neighbour; the owner is upwind where phi > 0): fc subtracted
    for (int j0 = l0; j0 < l1; j0 += kBatch)
    {
        int lf[kBatch];
        LinUpCell<NC> o[kBatch];
        double pl[kBatch], fx[kBatch], fy[kBatch], fz[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
        {
            const int j = min(j0 + u, l1 - 1);
            lf[u] = losort[j];
            linUpLoad<NC>(losortOwner[j], C, a, o[u]);
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
        {
            pl[u] = phi[lf[u]];
            fx[u] = Cf[3 * (size_t)lf[u]];
            fy[u] = Cf[3 * (size_t)lf[u] + 1];
            fz[u] = Cf[3 * (size_t)lf[u] + 2];
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
        {
            double fc[NC];
            linUpFace<NC>(pl[u], pl[u] > 0.0, self, o[u], fx[u], fy[u], fz[u], fc);
#pragma unroll
            for (int k = 0; k < NC; ++k)
                if (j0 + u < l1) s[k] -= fc[k];
        }
    }
    // the owned run (this cell upwind where phi > 0): fc added
    for (int g0 = f0; g0 < f1; g0 += kBatch)
    {
        int of[kBatch];
        LinUpCell<NC> o[kBatch];
        double po[kBatch], fx[kBatch], fy[kBatch], fz[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
        {
            of[u] = min(g0 + u, f1 - 1);
            po[u] = phi[of[u]];
            fx[u] = Cf[3 * (size_t)of[u]];
            fy[u] = Cf[3 * (size_t)of[u] + 1];
            fz[u] = Cf[3 * (size_t)of[u] + 2];
            linUpLoad<NC>(neighbour[of[u]], C, a, o[u]);
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
        {
            double fc[NC];
            linUpFace<NC>(po[u], !(po[u] > 0.0), self, o[u], fx[u], fy[u], fz[u], fc);
#pragma unroll
            for (int k = 0; k < NC; ++k)
                if (g0 + u < f1) s[k] += fc[k];
        }
    }
#pragma unroll
    for (int k = 0; k < NC; ++k) a.corr[k][c] = s[k];
}

// nuEff*dev2(T(grad(U))) per cell (linearViscousStress_cpp.cu:75-79; device_divdevreff.cu sigmaKernel):
// tensors component-major t[(i*3 + j)*n + c], gradU_ij = d(U_j)/d(x_i); sigma_ij = nu*(gradU_ji -
// (2/3)*tr on the diagonal)
__global__ __launch_bounds__(kTpb) void dev2tK(int n, const double* __restrict__ gradU, const double* __restrict__ nu,
                                               double* __restrict__ sigma)
{
    const int c = xcdBlock() * kTpb + threadIdx.x;
    if (c >= n) return;
    double g[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) g[q] = gradU[q * (size_t)n + c];
    const double k = (2.0 / 3.0) * (g[0] + g[4] + g[8]);
    const double v = nu[c];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
        {
            const double t = g[j * 3 + i];
            sigma[(i * 3 + j) * (size_t)n + c] = v * (i == j ? t - k : t);
        }
}

// V*fvc::div of a tensor field (fvc.cu:867-905 before the division by V; device_divdevreff.cu
// tensorDivKernel): per cell, its internal faces in face order -- the losort run subtracted, the owned
// run added -- with T_f = w*T_own + (1-w)*T_nei and (Sf & T_f)_j = Sx*T_xj + Sy*T_yj + Sz*T_zj, then its
// boundary faces not marked empty, Sf_b & T_b added. The cell's own tensor stays in registers; the
// other cell's tensor, w and Sf of up to kBatch faces are loaded together (clamped indices).
__device__ __forceinline__ void tensorFaceFlux(double wf, bool selfOwns, const double (&ts)[9],
                                               const double (&to)[9], double sx, double sy, double sz,
                                               double (&r)[3])
{
    double tf[9];
    const double wn = 1.0 - wf;
#pragma unroll
    for (int q = 0; q < 9; ++q) tf[q] = selfOwns ? wf * ts[q] + wn * to[q] : wf * to[q] + wn * ts[q];
#pragma unroll
    for (int j = 0; j < 3; ++j) r[j] = sx * tf[j] + sy * tf[3 + j] + sz * tf[6 + j];
}

// a cell's tensor: stored (Tc component-major), or nu*dev2(T(gradU)) formed from the gradient with
// dev2tK's arithmetic, so both give the same bits
template <bool FromGrad>
__device__ __forceinline__ void loadTensor(int n, int cell, const double* __restrict__ T,
                                           const double* __restrict__ nu, double (&t)[9])
{
    if (!FromGrad)
    {
#pragma unroll
        for (int q = 0; q < 9; ++q) t[q] = T[q * (size_t)n + cell];
        return;
    }
    double g[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) g[q] = T[q * (size_t)n + cell];
    const double k = (2.0 / 3.0) * (g[0] + g[4] + g[8]);
    const double v = nu[cell];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) t[i * 3 + j] = v * (i == j ? g[j * 3 + i] - k : g[j * 3 + i]);
}

template <bool Slab, bool FromGrad>
__global__ __launch_bounds__(kTpb) void tensorDivK(int n, SlabPart part, int nF, int nB,
                                                   const label* __restrict__ ownerStart,
                                                   const label* __restrict__ losortStart,
                                                   const label* __restrict__ losort,
                                                   const label* __restrict__ losortOwner,
                                                   const label* __restrict__ neighbour,
                                                   const label* __restrict__ bndStart,
                                                   const label* __restrict__ bndFaces,
                                                   const int* __restrict__ isEmpty, const double* __restrict__ w,
                                                   const double* __restrict__ Sf, const double* __restrict__ Tc,
                                                   const double* __restrict__ nu, const double* __restrict__ Tb,
                                                   double* __restrict__ dX,
                                                   double* __restrict__ dY, double* __restrict__ dZ)
{
    const int c = gatherCell<Slab>(n, part);
    if (c < 0) return;
    const int l0 = losortStart[c], l1 = losortStart[c + 1];
    const int f0 = ownerStart[c], f1 = ownerStart[c + 1];
    const int b0 = bndStart[c], b1 = bndStart[c + 1];
    double ts[9];
    loadTensor<FromGrad>(n, c, Tc, nu, ts);
    double d[3] = {0.0, 0.0, 0.0};
    for (int j0 = l0; j0 < l1; j0 += kBatch)
    {
        int lf[kBatch];
        double to[kBatch][9], wf[kBatch], sx[kBatch], sy[kBatch], sz[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
        {
            const int j = min(j0 + u, l1 - 1);
            lf[u] = losort[j];
            loadTensor<FromGrad>(n, losortOwner[j], Tc, nu, to[u]);
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
        {
            wf[u] = w[lf[u]];
            sx[u] = Sf[3 * (size_t)lf[u]];
            sy[u] = Sf[3 * (size_t)lf[u] + 1];
            sz[u] = Sf[3 * (size_t)lf[u] + 2];
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
        {
            double r[3];
            tensorFaceFlux(wf[u], false, ts, to[u], sx[u], sy[u], sz[u], r);
#pragma unroll
            for (int k = 0; k < 3; ++k)
                if (j0 + u < l1) d[k] = d[k] - r[k];
        }
    }
    for (int g0 = f0; g0 < f1; g0 += kBatch)
    {
        double to[kBatch][9], wf[kBatch], sx[kBatch], sy[kBatch], sz[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
        {
            const int f = min(g0 + u, f1 - 1);
            const int nb = neighbour[f];
            wf[u] = w[f];
            sx[u] = Sf[3 * (size_t)f];
            sy[u] = Sf[3 * (size_t)f + 1];
            sz[u] = Sf[3 * (size_t)f + 2];
            loadTensor<FromGrad>(n, nb, Tc, nu, to[u]);
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
        {
            double r[3];
            tensorFaceFlux(wf[u], true, ts, to[u], sx[u], sy[u], sz[u], r);
#pragma unroll
            for (int k = 0; k < 3; ++k)
                if (g0 + u < f1) d[k] += r[k];
        }
    }
    for (int j = b0; j < b1; ++j)
    {
        const int i = bndFaces[j];
        if (isEmpty && isEmpty[i]) continue;
        const double* s = Sf + 3 * ((size_t)nF + i);
#pragma unroll
        for (int k = 0; k < 3; ++k)
            d[k] += s[0] * Tb[(0 * 3 + k) * (size_t)nB + i] + s[1] * Tb[(1 * 3 + k) * (size_t)nB + i] +
                    s[2] * Tb[(2 * 3 + k) * (size_t)nB + i];
    }
    dX[c] = d[0];
    dY[c] = d[1];
    dZ[c] = d[2];
}

// per element: out = max(|a|, |b|, |c|) (cmptMax(cmptMag)), or min(a, b, c) (cmptMin); cf_types.cuh:199-202
template <bool Min>
__global__ __launch_bounds__(kTpb) void cmpt3K(int n, const double* __restrict__ x, const double* __restrict__ y,
                                               const double* __restrict__ z, double* __restrict__ out)
{
    const int i = xcdBlock() * kTpb + threadIdx.x;
    if (i >= n) return;
    out[i] = Min ? fmin(fmin(x[i], y[i]), z[i]) : fmax(fmax(fabs(x[i]), fabs(y[i])), fabs(z[i]));
}

// GeometricField::relax (pEqn.cu:422-433; pEqn_cpp.cu:295-300): p = pPrev + alpha*(p - pPrev)
__global__ __launch_bounds__(kTpb) void relaxFieldK(int n, double* __restrict__ p, const double* __restrict__ pPrev,
                                                    double alpha)
{
    const int i = xcdBlock() * kTpb + threadIdx.x;
    if (i >= n) return;
    const double q = pPrev[i];
    p[i] = q + alpha * (p[i] - q);
}

// constrainHbyA (pEqn.cu constrainHbyAKernel; pEqn_cpp.cu:55-72): U's boundary value on a face whose U
// is not assignable, the owner cell's HbyA elsewhere
__global__ __launch_bounds__(kTpb) void constrainHbyAK(int nB, const label* __restrict__ bFaceCell,
                                                       const int* __restrict__ takeU, const double* __restrict__ Ub,
                                                       const double* __restrict__ HbyA, double* __restrict__ HbyAb)
{
    const int i = xcdBlock() * kTpb + threadIdx.x;
    if (i >= nB) return;
    HbyAb[i] = takeU[i] ? Ub[i] : HbyA[bFaceCell[i]];
}

// flux of a cell vector field on the internal faces: (w U_own + (1-w) U_nei) . Sf (fvc.cu:545-563)
__global__ __launch_bounds__(kTpb) void vectorFluxK(int nF, const label* __restrict__ owner,
                                                    const label* __restrict__ neighbour,
                                                    const double* __restrict__ w, const double* __restrict__ Sf,
                                                    const double* __restrict__ Ux, const double* __restrict__ Uy,
                                                    const double* __restrict__ Uz, double* __restrict__ phi)
{
    const int f = xcdBlock() * kTpb + threadIdx.x;
    if (f >= nF) return;
    const int o = owner[f], n = neighbour[f];
    const double wf = w[f];
    const double ux = wf * Ux[o] + (1.0 - wf) * Ux[n];
    const double uy = wf * Uy[o] + (1.0 - wf) * Uy[n];
    const double uz = wf * Uz[o] + (1.0 - wf) * Uz[n];
    const double* s = Sf + 3 * (size_t)f;
    phi[f] = ux * s[0] + uy * s[1] + uz * s[2];
}

// ...and on the boundary faces: Ub . Sf_b, zero on an empty face (fvc.cu:564-580)
__global__ __launch_bounds__(kTpb) void boundaryFluxK(int nB, const double* __restrict__ SfB,
                                                      const double* __restrict__ uxb, const double* __restrict__ uyb,
                                                      const double* __restrict__ uzb, const int* __restrict__ isEmpty,
                                                      double* __restrict__ phiB)
{
    const int i = xcdBlock() * kTpb + threadIdx.x;
    if (i >= nB) return;
    const double* s = SfB + 3 * (size_t)i;
    phiB[i] = (isEmpty && isEmpty[i]) ? 0.0 : uxb[i] * s[0] + uyb[i] * s[1] + uzb[i] * s[2];
}

// adjustPhi's sums (pEqn_cpp.cu:85-105): per workgroup, sum|phi| over its chunk of the internal faces (eight
// loads in flight per thread), and massIn, fixed outflow and adjustable outflow over its chunk of the
// boundary faces (four partial slots)
__global__ __launch_bounds__(kReduceBlock) void adjustSumsK(int nF, const double* __restrict__ phiInt, int nB,
                                                            const int* __restrict__ adjustable,
                                                            const double* __restrict__ phiB, double* pTot,
                                                            double* pIn, double* pFix, double* pAdj)
{
    __shared__ double lds[kReduceBlock / 64];
    const int wb = xcdBlock();
    double t = 0.0, in = 0.0, fix = 0.0, adj = 0.0;
    {
        const int ch = chunkRows(nF, gridDim.x), b = wb * ch, e = b + ch < nF ? b + ch : nF;
        int i = b + (int)threadIdx.x;
        for (; i + 7 * kReduceBlock < e; i += 8 * kReduceBlock)
        {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = phiInt[i + u * kReduceBlock];
#pragma unroll
            for (int u = 0; u < 8; ++u) t += fabs(v[u]);
        }
        for (; i < e; i += kReduceBlock) t += fabs(phiInt[i]);
    }
    {
        const int ch = chunkRows(nB, gridDim.x), b = wb * ch, e = b + ch < nB ? b + ch : nB;
        for (int i = b + (int)threadIdx.x; i < e; i += kReduceBlock)
        {
            const double v = phiB[i];
            const bool a = adjustable[i] != 0;
            const bool out = !(v < 0.0);   // the host's else-branch (a NaN included)
            in += out ? 0.0 : -v;
            fix += (out && !a) ? v : 0.0;
            adj += (out && a) ? v : 0.0;
        }
    }
    const double st = blockSum(t, lds);
    const double si = blockSum(in, lds);
    const double sf = blockSum(fix, lds);
    const double sa = blockSum(adj, lds);
    if (threadIdx.x == 0)
    {
        pTot[blockIdx.x] = st;
        pIn[blockIdx.x] = si;
        pFix[blockIdx.x] = sf;
        pAdj[blockIdx.x] = sa;
    }
}

// the decision (pEqn_cpp.cu:107-123; device_simple.cu:282-346): out = {massCorr, refused, closedVolume}
__device__ void adjustDecide(double tot, double in, double fix, double adj, double* out)
{
    const double vsmall = 1e-300, small = 1e-15;
    const double totalFlux = tot;
    double mc = 1.0, refused = 0.0;
    const double magAdj = fabs(adj);
    if (magAdj > vsmall && magAdj / totalFlux > small)
        mc = (in - fix) / adj;
    else if (fabs(fix - in) / totalFlux > 1e-8)
        refused = 1.0;
    out[0] = mc;
    out[1] = refused;
    out[2] = (fabs(in) / totalFlux < small && fabs(fix) / totalFlux < small && fabs(adj) / totalFlux < small) ? 1.0
                                                                                                              : 0.0;
}

__global__ __launch_bounds__(kFinishBlock) void adjustFinishK(int nParts, const double* __restrict__ pTot,
                                                              const double* __restrict__ pIn,
                                                              const double* __restrict__ pFix,
                                                              const double* __restrict__ pAdj, double* out)
{
    __shared__ double lds[kFinishBlock / 64];
    const double t = blockSum<kFinishBlock>(finishShare(nParts, pTot), lds);
    const double in = blockSum<kFinishBlock>(finishShare(nParts, pIn), lds);
    const double fix = blockSum<kFinishBlock>(finishShare(nParts, pFix), lds);
    const double adj = blockSum<kFinishBlock>(finishShare(nParts, pAdj), lds);
    if (threadIdx.x == 0) adjustDecide(1e-300 + t, in, fix, adj, out);
}

// the same sums in the host loops' order on one thread (the sequential-reduction mode)
__global__ void adjustSeqK(int nF, const double* __restrict__ phiInt, int nB, const int* __restrict__ adjustable,
                           const double* __restrict__ phiB, double* out)
{
    if (threadIdx.x || blockIdx.x) return;
    double t = 1e-300, in = 0.0, fix = 0.0, adj = 0.0;
    for (int f = 0; f < nF; ++f) t += fabs(phiInt[f]);
    for (int i = 0; i < nB; ++i)
    {
        const double v = phiB[i];
        if (v < 0.0)
            in -= v;
        else if (!adjustable[i])
            fix += v;
        else
            adj += v;
    }
    adjustDecide(t, in, fix, adj, out);
}

// scale the adjustable outflow (pEqn_cpp.cu:125-131); nothing when the correction was refused
__global__ __launch_bounds__(kTpb) void adjustScaleK(int nB, const int* __restrict__ adjustable,
                                                     const double* __restrict__ decision, double* __restrict__ phiB)
{
    const int i = xcdBlock() * kTpb + threadIdx.x;
    if (i >= nB) return;
    const double mc = decision[0];
    if (decision[1] == 0.0 && adjustable[i] && phiB[i] > 0.0) phiB[i] *= mc;
}

// zeros (a kernel, not a memset, so the call can sit in a captured graph)
__global__ __launch_bounds__(kTpb) void zeroK(int n, double* __restrict__ x)
{
    const int i = xcdBlock() * kTpb + threadIdx.x;
    if (i < n) x[i] = 0.0;
}

// The Gauss linear gradient of NC cell fields that share their patches' types (fvc.cu:67-121), each
// component's arithmetic the host loop's: per cell its faces in face order -- the losort run (-Sf*pf), the owned run
// (+Sf*pf), pf = w*(P - N) + N -- then its boundary faces in order (empty: none; fixedValue: the component's
// bc value; zeroGradient: the cell value), over V. Each face's weight, area vector and neighbour index are
// loaded once for all NC components. Writes each component's gradient xyz per cell (g3) and, when g9 is
// given, the packed tensor g9[(i*3 + k)*n + c] = d(psi_k)/dx_i.
struct Grad3Args
{
    const double* psi[3];
    const double* bcValue[3];
    double* g3[3];
    double* g9;
};

template <bool Slab, int NC>
__global__ __launch_bounds__(kTpb) void gradGauss3K(int n, SlabPart part, int nF, const label* __restrict__ ownerStart,
                                                    const label* __restrict__ losortStart,
                                                    const label* __restrict__ losort,
                                                    const label* __restrict__ losortOwner,
                                                    const label* __restrict__ neighbour,
                                                    const label* __restrict__ bndStart,
                                                    const label* __restrict__ bndFaces,
                                                    const int* __restrict__ bcType, const double* __restrict__ Sf,
                                                    const double* __restrict__ w, const double* __restrict__ V,
                                                    Grad3Args a)
{
    const int c = gatherCell<Slab>(n, part);
    if (c < 0) return;
    double pc[NC], gx[NC], gy[NC], gz[NC];
#pragma unroll
    for (int j = 0; j < NC; ++j)
    {
        pc[j] = a.psi[j][c];
        gx[j] = 0.0;
        gy[j] = 0.0;
        gz[j] = 0.0;
    }
    for (int k = losortStart[c]; k < losortStart[c + 1]; ++k)
    {
        const int f = losort[k], o = losortOwner[k];
        const double wf = w[f], sx = Sf[3 * (size_t)f], sy = Sf[3 * (size_t)f + 1], sz = Sf[3 * (size_t)f + 2];
#pragma unroll
        for (int j = 0; j < NC; ++j)
        {
            const double pf = wf * (a.psi[j][o] - pc[j]) + pc[j];
            gx[j] = gx[j] - sx * pf;
            gy[j] = gy[j] - sy * pf;
            gz[j] = gz[j] - sz * pf;
        }
    }
    for (int f = ownerStart[c]; f < ownerStart[c + 1]; ++f)
    {
        const int nb = neighbour[f];
        const double wf = w[f], sx = Sf[3 * (size_t)f], sy = Sf[3 * (size_t)f + 1], sz = Sf[3 * (size_t)f + 2];
#pragma unroll
        for (int j = 0; j < NC; ++j)
        {
            const double pn = a.psi[j][nb];
            const double pf = wf * (pc[j] - pn) + pn;
            gx[j] = gx[j] + sx * pf;
            gy[j] = gy[j] + sy * pf;
            gz[j] = gz[j] + sz * pf;
        }
    }
    for (int k = bndStart[c]; k < bndStart[c + 1]; ++k)
    {
        const int i = bndFaces[k];
        const int t = bcType[i];
        if (t == 2) continue;
        const size_t fb = 3 * ((size_t)nF + i);
        const double sx = Sf[fb], sy = Sf[fb + 1], sz = Sf[fb + 2];
#pragma unroll
        for (int j = 0; j < NC; ++j)
        {
            const double pv = t == 1 ? a.bcValue[j][i] : pc[j];
            gx[j] = gx[j] + sx * pv;
            gy[j] = gy[j] + sy * pv;
            gz[j] = gz[j] + sz * pv;
        }
    }
    const double v = V[c];
#pragma unroll
    for (int j = 0; j < NC; ++j)
    {
        const double x = gx[j] / v, y = gy[j] / v, z = gz[j] / v;
        if (a.g3[j])
        {
            a.g3[j][3 * (size_t)c] = x;
            a.g3[j][3 * (size_t)c + 1] = y;
            a.g3[j][3 * (size_t)c + 2] = z;
        }
        if (a.g9)
        {
            a.g9[(size_t)(0 * 3 + j) * n + c] = x;
            a.g9[(size_t)(1 * 3 + j) * n + c] = y;
            a.g9[(size_t)(2 * 3 + j) * n + c] = z;
        }
    }
}

} // namespace

template <int NC>
static void launchGrad(Ctx& c, const Mesh& m, const int* bcType, const double* Sf, const double* w, const double* V,
                       const Grad3Args& a)
{
    const SlabPart part = gatherPart(m);
    if (part.plane)
        gradGauss3K<true, NC><<<gatherBlocks(m.nCells, part.plane), kTpb, 0, c.stream>>>(
            m.nCells, part, m.nFaces, m.ownerStart.p, m.losortStart.p, m.losort.p, m.losortOwner.p, m.neighbour.p,
            m.bndStart.p, m.bndFaces.p, bcType, Sf, w, V, a);
    else
        gradGauss3K<false, NC><<<blocks(m.nCells), kTpb, 0, c.stream>>>(
            m.nCells, part, m.nFaces, m.ownerStart.p, m.losortStart.p, m.losort.p, m.losortOwner.p, m.neighbour.p,
            m.bndStart.p, m.bndFaces.p, bcType, Sf, w, V, a);
    hipCheck(hipGetLastError(), "grad gauss");
}

void launchGradGauss3(Ctx& c, const Mesh& m, const int* bcType, const double* const* bcValue, const double* Sf,
                      const double* w, const double* V, const double* const* psi, double* const* g3, double* g9)
{
    if (!m.nCells) return;
    Grad3Args a{};
    for (int j = 0; j < 3; ++j)
    {
        a.psi[j] = psi[j];
        a.bcValue[j] = bcValue[j];
        a.g3[j] = g3[j];
    }
    a.g9 = g9;
    launchGrad<3>(c, m, bcType, Sf, w, V, a);
}

void launchGradGauss(Ctx& c, const Mesh& m, const int* bcType, const double* bcValue, const double* Sf,
                     const double* w, const double* V, const double* psi, double* grad)
{
    if (!m.nCells) return;
    Grad3Args a{};
    a.psi[0] = psi;
    a.bcValue[0] = bcValue;
    a.g3[0] = grad;
    launchGrad<1>(c, m, bcType, Sf, w, V, a);
}

void launchZero(Ctx& c, int n, double* x)
{
    if (n <= 0) return;
    zeroK<<<blocks(n), kTpb, 0, c.stream>>>(n, x);
    hipCheck(hipGetLastError(), "zero");
}

void launchFoldBoundaryDiag(Ctx& c, const Mesh& m, const double* icX, const double* icY, const double* icZ,
                            const double* diag, double* D)
{
    if (m.nCells == 0) return;
    foldBoundaryDiagK<<<blocks(m.nCells), kTpb, 0, c.stream>>>(m.nCells, m.bndStart.p, m.bndFaces.p, icX, icY, icZ,
                                                                diag, D);
    hipCheck(hipGetLastError(), "fold boundary diag");
}

void launchReciprocalV(Ctx& c, const Mesh& m, const double* D, const double* V, double* rAU)
{
    if (m.nCells == 0) return;
    reciprocalVK<<<blocks(m.nCells), kTpb, 0, c.stream>>>(m.nCells, D, V, rAU);
    hipCheck(hipGetLastError(), "reciprocal V");
}

void launchBoundaryDiagTerm(Ctx& c, int nB, const double* icX, const double* icY, const double* icZ, int k,
                            double* out)
{
    if (nB == 0) return;
    boundaryDiagTermK<<<blocks(nB), kTpb, 0, c.stream>>>(nB, icX, icY, icZ, k, out);
    hipCheck(hipGetLastError(), "boundary diag term");
}

void launchMatrixH(Ctx& c, const Mesh& m, const double* upper, const double* lower, const double* psi,
                   const double* source, const double* bdDiag, const double* bdSrc, const double* V,
                   const double* rAU, double* out)
{
    if (m.nCells == 0) return;
    const SlabPart part = gatherPart(m);
    if (part.plane)
        matrixHK<true><<<gatherBlocks(m.nCells, part.plane), kTpb, 0, c.stream>>>(
            m.nCells, part, m.ownerStart.p, m.losortStart.p, m.losort.p, m.losortOwner.p, m.neighbour.p, m.bndStart.p,
            m.bndFaces.p, upper, lower, psi, source, bdDiag, bdSrc, V, rAU, out);
    else
        matrixHK<false><<<blocks(m.nCells), kTpb, 0, c.stream>>>(
            m.nCells, part, m.ownerStart.p, m.losortStart.p, m.losort.p, m.losortOwner.p, m.neighbour.p, m.bndStart.p,
            m.bndFaces.p, upper, lower, psi, source, bdDiag, bdSrc, V, rAU, out);
    hipCheck(hipGetLastError(), "matrix H");
}

void launchMatrixH3(Ctx& c, const Mesh& m, const double* upper, const double* lower, const double* const* psi,
                    const double* const* source, const double* const* bdDiag, const double* const* bdSrc,
                    const double* V, const double* rAU, const int* valid, double* const* out)
{
    if (m.nCells == 0) return;
    Vec3Args a{};
    for (int k = 0; k < 3; ++k)
    {
        a.valid[k] = valid ? (valid[k] != 0) : 1;
        // a knocked-out component reads the first valid one's inputs (in bounds) and writes zeros
        const int s = a.valid[k] ? k : (a.valid[0] ? 0 : (a.valid[1] ? 1 : 2));
        a.psi[k] = psi[s];
        a.source[k] = source[s];
        a.bdDiag[k] = bdDiag[s];
        a.bdSrc[k] = bdSrc[s];
        a.out[k] = out[k];
    }
    const SlabPart part = gatherPart(m);
    if (part.plane)
        matrixH3K<true><<<gatherBlocks(m.nCells, part.plane), kTpb, 0, c.stream>>>(
            m.nCells, part, m.nFaces, m.ownerStart.p, m.losortStart.p, m.losort.p, m.losortOwner.p, m.neighbour.p,
            m.bndStart.p, m.bndFaces.p, upper, lower, a, V, rAU);
    else
        matrixH3K<false><<<blocks(m.nCells), kTpb, 0, c.stream>>>(
            m.nCells, part, m.nFaces, m.ownerStart.p, m.losortStart.p, m.losort.p, m.losortOwner.p, m.neighbour.p,
            m.bndStart.p, m.bndFaces.p, upper, lower, a, V, rAU);
    hipCheck(hipGetLastError(), "matrix H (three components)");
}

void launchRelaxDiag(Ctx& c, const Mesh& m, const double* upper, const double* lower, const double* diag,
                     const double* icBnd, const double* icMag, const double* icMin, const double* cycSumOff,
                     double alpha, int nComp, const double* const* psi, double* const* source, double* relaxed,
                     double* delta)
{
    if (m.nCells == 0) return;
    RelaxArgs a{};
    a.icBnd = icBnd;
    a.icMag = icMag;
    a.icMin = icMin;
    a.cycSumOff = cycSumOff;
    a.alpha = alpha;
    a.nComp = nComp;
    for (int k = 0; k < nComp; ++k)
    {
        a.psi[k] = psi[k];
        a.source[k] = source[k];
    }
    const SlabPart part = gatherPart(m);
    if (part.plane)
        relaxK<true><<<gatherBlocks(m.nCells, part.plane), kTpb, 0, c.stream>>>(
            m.nCells, part, m.nFaces, m.ownerStart.p, m.losortStart.p, m.losort.p, m.bndStart.p, m.bndFaces.p, upper,
            lower, diag, a, relaxed, delta);
    else
        relaxK<false><<<blocks(m.nCells), kTpb, 0, c.stream>>>(m.nCells, part, m.nFaces, m.ownerStart.p,
                                                               m.losortStart.p, m.losort.p, m.bndStart.p,
                                                               m.bndFaces.p, upper, lower, diag, a, relaxed, delta);
    hipCheck(hipGetLastError(), "relax diag");
}

template <int NC>
static void launchLinUp(Ctx& c, const Mesh& m, const double* phi, const double* Cf, const double* C,
                        const LinUpArgs& a)
{
    const SlabPart part = gatherPart(m);
    if (part.plane)
        linearUpwindCorrK<true, NC><<<gatherBlocks(m.nCells, part.plane, kLinUpSplit), kTpb, 0, c.stream>>>(
            m.nCells, part, m.nFaces, m.ownerStart.p, m.losortStart.p, m.losort.p, m.losortOwner.p, m.neighbour.p, phi,
            Cf, C, a);
    else
        linearUpwindCorrK<false, NC><<<blocks(m.nCells), kTpb, 0, c.stream>>>(
            m.nCells, part, m.nFaces, m.ownerStart.p, m.losortStart.p, m.losort.p, m.losortOwner.p, m.neighbour.p, phi,
            Cf, C, a);
}

void launchLinearUpwindCorr(Ctx& c, const Mesh& m, const double* phi, const double* Cf, const double* C, int nComp,
                            const double* const* grad, double* const* corr)
{
    if (m.nCells == 0) return;
    LinUpArgs a{};
    for (int k = 0; k < nComp; ++k)
    {
        a.grad[k] = grad[k];
        a.corr[k] = corr[k];
    }
    if (nComp == 3)
        launchLinUp<3>(c, m, phi, Cf, C, a);
    else if (nComp == 2)
        launchLinUp<2>(c, m, phi, Cf, C, a);
    else
        launchLinUp<1>(c, m, phi, Cf, C, a);
    hipCheck(hipGetLastError(), "linearUpwind correction");
}

void launchScaledDev2T(Ctx& c, int n, const double* gradU, const double* nu, double* sigma)
{
    if (n <= 0) return;
    dev2tK<<<blocks(n), kTpb, 0, c.stream>>>(n, gradU, nu, sigma);
    hipCheck(hipGetLastError(), "nu*dev2(T(gradU))");
}

template <bool FromGrad>
static void launchTensorDivT(Ctx& c, const Mesh& m, const double* w, const double* Sf, const int* isEmpty,
                             const double* Tc, const double* nu, const double* Tb, double* dX, double* dY, double* dZ)
{
    const SlabPart part = gatherPart(m);
    if (part.plane)
        tensorDivK<true, FromGrad><<<gatherBlocks(m.nCells, part.plane), kTpb, 0, c.stream>>>(
            m.nCells, part, m.nFaces, m.nBFaces, m.ownerStart.p, m.losortStart.p, m.losort.p, m.losortOwner.p,
            m.neighbour.p, m.bndStart.p, m.bndFaces.p, isEmpty, w, Sf, Tc, nu, Tb, dX, dY, dZ);
    else
        tensorDivK<false, FromGrad><<<blocks(m.nCells), kTpb, 0, c.stream>>>(
            m.nCells, part, m.nFaces, m.nBFaces, m.ownerStart.p, m.losortStart.p, m.losort.p, m.losortOwner.p,
            m.neighbour.p, m.bndStart.p, m.bndFaces.p, isEmpty, w, Sf, Tc, nu, Tb, dX, dY, dZ);
}

// nu == nullptr: Tc holds the tensors; else Tc holds grad(U) and each cell's nu*dev2(T(grad U)) is formed
// where it is read
void launchTensorDiv(Ctx& c, const Mesh& m, const double* w, const double* Sf, const int* isEmpty, const double* Tc,
                     const double* nu, const double* Tb, double* dX, double* dY, double* dZ)
{
    if (m.nCells == 0) return;
    if (nu)
        launchTensorDivT<true>(c, m, w, Sf, isEmpty, Tc, nu, Tb, dX, dY, dZ);
    else
        launchTensorDivT<false>(c, m, w, Sf, isEmpty, Tc, nu, Tb, dX, dY, dZ);
    hipCheck(hipGetLastError(), "tensor divergence");
}

void launchCmpt3(Ctx& c, int n, const double* x, const double* y, const double* z, bool min, double* out)
{
    if (n <= 0) return;
    if (min)
        cmpt3K<true><<<blocks(n), kTpb, 0, c.stream>>>(n, x, y, z, out);
    else
        cmpt3K<false><<<blocks(n), kTpb, 0, c.stream>>>(n, x, y, z, out);
    hipCheck(hipGetLastError(), "component max/min");
}

void launchRelaxField(Ctx& c, int n, double* p, const double* pPrev, double alpha)
{
    if (n <= 0) return;
    relaxFieldK<<<blocks(n), kTpb, 0, c.stream>>>(n, p, pPrev, alpha);
    hipCheck(hipGetLastError(), "relax field");
}

void launchConstrainHbyA(Ctx& c, const Mesh& m, const int* takeU, const double* Ub, const double* HbyA,
                         double* HbyAb)
{
    if (m.nBFaces == 0) return;
    constrainHbyAK<<<blocks(m.nBFaces), kTpb, 0, c.stream>>>(m.nBFaces, m.bFaceCell.p, takeU, Ub, HbyA, HbyAb);
    hipCheck(hipGetLastError(), "constrain HbyA");
}

void launchVectorFlux(Ctx& c, const Mesh& m, const double* w, const double* Sf, const double* Ux, const double* Uy,
                      const double* Uz, double* phi)
{
    if (m.nFaces == 0) return;
    vectorFluxK<<<blocks(m.nFaces), kTpb, 0, c.stream>>>(m.nFaces, m.owner.p, m.neighbour.p, w, Sf, Ux, Uy, Uz,
                                                          phi);
    hipCheck(hipGetLastError(), "vector flux");
}

void launchBoundaryFlux(Ctx& c, const Mesh& m, const double* SfB, const double* uxb, const double* uyb,
                        const double* uzb, const int* isEmpty, double* phiB)
{
    if (m.nBFaces == 0) return;
    boundaryFluxK<<<blocks(m.nBFaces), kTpb, 0, c.stream>>>(m.nBFaces, SfB, uxb, uyb, uzb, isEmpty, phiB);
    hipCheck(hipGetLastError(), "boundary flux");
}

void launchAdjustPhi(Ctx& c, int nB, const int* adjustable, double* phiB, int nF, const double* phiInt,
                     double* decision)
{
    if (!phiInt) nF = 0;
    if (c.sequentialReduce)
        adjustSeqK<<<1, 64, 0, c.stream>>>(nF, phiInt, nB, adjustable, phiB, decision);
    else
    {
        const int g = reduceGridRows(nF > nB ? nF : nB, 8);
        double* pa = c.partials.p;
        adjustSumsK<<<g, kReduceBlock, 0, c.stream>>>(nF, phiInt, nB, adjustable, phiB, pa, pa + kMaxReduceBlocks,
                                                      pa + 2 * (size_t)kMaxReduceBlocks,
                                                      pa + 3 * (size_t)kMaxReduceBlocks);
        adjustFinishK<<<1, kFinishBlock, 0, c.stream>>>(g, pa, pa + kMaxReduceBlocks, pa + 2 * (size_t)kMaxReduceBlocks,
                                                        pa + 3 * (size_t)kMaxReduceBlocks, decision);
    }
    if (nB) adjustScaleK<<<blocks(nB), kTpb, 0, c.stream>>>(nB, adjustable, decision, phiB);
    hipCheck(hipGetLastError(), "adjust phi");
}

} // namespace mfv
