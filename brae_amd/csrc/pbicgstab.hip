// PBiCGStab for the asymmetric matrices of the segregated solves (the momentum components, and the
// turbulence scalars), device-resident: the recurrence of src/OpenFOAM/matrices/pbicgstab.cu:6-150 (the
// solver solveVector hands each component to, solve_vector.cuh:27-116), with its fold of the boundary
// coefficients, OpenFOAM's normFactor and the 1e-300 breakdown guards.
//
// Preconditioners:
//   DILU      the reference's: the reciprocal diagonal rD[c] = diag - sum over lower faces (ascending) of
//             upper*lower/rD[lower neighbour], the forward substitution over lower faces ascending and the
//             backward one over owned faces descending. Each is a sequential walk; it runs on the mesh's
//             level schedule (one launch per dependency level, one cell per thread) in the host loop's
//             per-cell order, so with sequential reductions the solve is the oracle's to the bit.
//   diagonal  y = p / diag: one stream, for meshes whose level schedule is too deep to walk fast.
//
// Scalars never leave the device: one-thread state kernels form alpha, beta, omega and the residuals
// after each reduction and set a stop flag that every later kernel of the solve checks; the host polls
// the flag every checkEvery iterations.
#include "amg.hpp"
#include "reduce.hpp"

#include <mutex>

namespace mfv {

// The PBiCGStab's device buffers, kept on the mesh between solves (resized when the mesh's size changes).
struct BicgWork
{
    DBuf<double> diagC, b, rD, r, r0, p, y, Ay, s, z, t, st;
    DBuf<int> fl;
    void size(int n);
};

// The device addresses one solve works in: the mesh's BicgWork, or the static workspace below.
struct BicgPtrs
{
    double *diagC, *b, *rD, *r, *r0, *p, *y, *Ay, *s, *z, *t, *st;
    int* fl;
};

void BicgWorkDeleter::operator()(BicgWork* w) const { delete w; }

namespace {

constexpr int kTpb = 256;
inline int blocks(int n) { return (n + kTpb - 1) / kTpb; }

enum : int { B_RHO = 0, B_RHOOLD, B_ALPHA, B_OMEGA, B_BETA, B_NF, B_INIT, B_FINAL, B_TOL, B_RELTOL, B_T0, B_T1,
             B_T2, B_T3, B_T4, B_SUMPSI, B_COUNT = 24 };
enum : int { BF_ITER = 0, BF_STOP, BF_CONV, BF_MAXIT, BF_MINIT, BF_EXIT, BF_COUNT = 8 };

__device__ __forceinline__ bool bicgConverged(const double* st, double fr)
{
    return fr < st[B_TOL] || (st[B_RELTOL] > 0.0 && fr < st[B_RELTOL] * st[B_INIT]);
}

__global__ void bicgSetupK(double* st, int* fl, double tol, double relTol, int maxIter, int minIter)
{
    if (threadIdx.x || blockIdx.x) return;
    st[B_TOL] = tol;
    st[B_RELTOL] = relTol;
    st[B_RHO] = 0.0;
    st[B_ALPHA] = 0.0;
    st[B_OMEGA] = 0.0;
    fl[BF_ITER] = 0;
    fl[BF_STOP] = 0;
    fl[BF_EXIT] = 0;
    fl[BF_MAXIT] = maxIter;
    fl[BF_MINIT] = minIter;
}

// fold: diagC = diag + iC, b = source + bC over each cell's boundary faces ascending (pbicgstab.cu:24-31)
__global__ __launch_bounds__(kTpb) void bicgFoldK(int n, const label* __restrict__ bndStart,
                                                  const label* __restrict__ bndFaces, const double* __restrict__ diag,
                                                  const double* __restrict__ source, const double* __restrict__ iC,
                                                  const double* __restrict__ bC, double* __restrict__ diagC,
                                                  double* __restrict__ b)
{
    const int c = xcdBlock() * kTpb + threadIdx.x;
    if (c >= n) return;
    double d = diag[c], s = source[c];
    for (int k = bndStart[c]; k < bndStart[c + 1]; ++k)
    {
        const int i = bndFaces[k];
        d += iC[i];
        s += bC[i];
    }
    diagC[c] = d;
    b[c] = s;
}

// y = A x over a cell's faces in face order (the losort run ascending, then the owned run ascending), unless
// stopped. The first kAmulBatch faces of both runs are loaded together, each level of the gather (face,
// coefficient and neighbour, then x) at once: clamped indices keep every load in bounds and unconditional,
// and a face beyond its run is dropped from the sum, not added as zero; later faces follow in a loop. With
// Slab the cells follow the plane-slab partition of the mesh's band, so a band-back neighbour's x and lower
// coefficient were read by the same XCD one slab earlier and are still in its L2.
constexpr int kAmulBatch = 3;

template <bool Slab>
__global__ __launch_bounds__(kTpb) void bicgAmulK(int n, SlabPart part, int nF, const label* __restrict__ ownerStart,
                                                  const label* __restrict__ losortStart,
                                                  const label* __restrict__ losort,
                                                  const label* __restrict__ losortOwner,
                                                  const label* __restrict__ neighbour, const double* __restrict__ diag,
                                                  const double* __restrict__ upper, const double* __restrict__ lower,
                                                  const double* __restrict__ x, double* __restrict__ y,
                                                  const int* __restrict__ fl)
{
    if (fl[BF_STOP]) return;
    const int c = gatherCell<Slab>(n, part);
    if (c < 0) return;
    const int l0 = losortStart[c], l1 = losortStart[c + 1];
    const int f0 = ownerStart[c], f1 = ownerStart[c + 1];
    double s = diag[c] * x[c];
    double lv[kAmulBatch] = {}, lx[kAmulBatch] = {}, uv[kAmulBatch] = {}, ux[kAmulBatch] = {};
    if (nF > 0)   // uniform: a mesh with internal faces
    {
        const int last = nF - 1;
        int lf[kAmulBatch], lo[kAmulBatch], nb[kAmulBatch];
#pragma unroll
        for (int u = 0; u < kAmulBatch; ++u)
        {
            const int j = min(l0 + u, last), f = min(f0 + u, last);
            lf[u] = losort[j];
            lo[u] = losortOwner[j];
            nb[u] = neighbour[f];
            uv[u] = upper[f];
        }
#pragma unroll
        for (int u = 0; u < kAmulBatch; ++u)
        {
            lv[u] = lower[lf[u]];
            lx[u] = x[lo[u]];
            ux[u] = x[nb[u]];
        }
    }
#pragma unroll
    for (int u = 0; u < kAmulBatch; ++u)
        if (l0 + u < l1) s += lv[u] * lx[u];
    for (int k = l0 + kAmulBatch; k < l1; ++k) s += lower[losort[k]] * x[losortOwner[k]];
#pragma unroll
    for (int u = 0; u < kAmulBatch; ++u)
        if (f0 + u < f1) s += uv[u] * ux[u];
    for (int f = f0 + kAmulBatch; f < f1; ++f) s += upper[f] * x[neighbour[f]];
    y[c] = s;
}

__global__ __launch_bounds__(kTpb) void bicgSubK(int n, const double* __restrict__ b, const double* __restrict__ ax,
                                                 double* __restrict__ r, double* __restrict__ r0)
{
    const int c = blockIdx.x * kTpb + threadIdx.x;
    if (c >= n) return;
    const double v = b[c] - ax[c];
    r[c] = v;
    r0[c] = v;
}

// ---- DILU on the level schedule ----
__global__ __launch_bounds__(kTpb) void diluFactorK(int cnt, const int* __restrict__ cells,
                                                    const label* __restrict__ losortStart,
                                                    const label* __restrict__ losort,
                                                    const label* __restrict__ losortOwner,
                                                    const double* __restrict__ diag, const double* __restrict__ upper,
                                                    const double* __restrict__ lower, double* __restrict__ rD)
{
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= cnt) return;
    const int c = cells[i];
    double d = diag[c];
    for (int k = losortStart[c]; k < losortStart[c + 1]; ++k)
    {
        const int f = losort[k];
        d -= upper[f] * lower[f] / rD[losortOwner[k]];
    }
    rD[c] = d;
}

__global__ __launch_bounds__(kTpb) void invertK(int n, double* __restrict__ rD)
{
    const int c = blockIdx.x * kTpb + threadIdx.x;
    if (c < n) rD[c] = 1.0 / rD[c];
}

// forward: y[c] = rD[c]*r[c], then y[c] -= rD[c]*lower*y[lower neighbour] over lower faces ascending
__global__ __launch_bounds__(kTpb) void diluFwdK(int cnt, const int* __restrict__ cells,
                                                 const label* __restrict__ losortStart,
                                                 const label* __restrict__ losort,
                                                 const label* __restrict__ losortOwner, const double* __restrict__ rD,
                                                 const double* __restrict__ lower, const double* __restrict__ r,
                                                 double* __restrict__ y, const int* __restrict__ fl)
{
    if (fl[BF_STOP]) return;
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= cnt) return;
    const int c = cells[i];
    const double rd = rD[c];
    double v = rd * r[c];
    for (int k = losortStart[c]; k < losortStart[c + 1]; ++k) v -= rd * lower[losort[k]] * y[losortOwner[k]];
    y[c] = v;
}

// backward: y[c] -= rD[c]*upper*y[upper neighbour] over owned faces descending
__global__ __launch_bounds__(kTpb) void diluBwdK(int cnt, const int* __restrict__ cells,
                                                 const label* __restrict__ ownerStart,
                                                 const label* __restrict__ neighbour, const double* __restrict__ rD,
                                                 const double* __restrict__ upper, double* __restrict__ y,
                                                 const int* __restrict__ fl)
{
    if (fl[BF_STOP]) return;
    const int i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= cnt) return;
    const int c = cells[i];
    const double rd = rD[c];
    double v = y[c];
    for (int f = ownerStart[c + 1] - 1; f >= ownerStart[c]; --f) v -= rd * upper[f] * y[neighbour[f]];
    y[c] = v;
}

// diagonal preconditioner: y = rD*r with rD = 1/diag
__global__ __launch_bounds__(kTpb) void diagApplyK(int n, const double* __restrict__ rD, const double* __restrict__ r,
                                                   double* __restrict__ y, const int* __restrict__ fl)
{
    if (fl[BF_STOP]) return;
    const int c = blockIdx.x * kTpb + threadIdx.x;
    if (c < n) y[c] = rD[c] * r[c];
}

__global__ __launch_bounds__(kTpb) void diagInvK(int n, const double* __restrict__ d, double* __restrict__ rD)
{
    const int c = blockIdx.x * kTpb + threadIdx.x;
    if (c < n) rD[c] = 1.0 / d[c];
}

// ---- stop-aware reductions into a device scalar ----
__global__ __launch_bounds__(kReduceBlock) void dotStopK(int n, const double* __restrict__ a,
                                                         const double* __restrict__ b, double* __restrict__ partials,
                                                         const int* __restrict__ fl)
{
    if (fl[BF_STOP]) return;
    __shared__ double lds[kReduceBlock / 64];
    double v = 0.0;
    MFV_ROWS(n, i) v += a[i] * b[i];
    storePartial(blockSum(v, lds), partials);
}

__global__ __launch_bounds__(kReduceBlock) void dot2StopK(int n, const double* __restrict__ t,
                                                          const double* __restrict__ s, double* __restrict__ pa,
                                                          double* __restrict__ pb, const int* __restrict__ fl)
{
    if (fl[BF_STOP]) return;
    __shared__ double lds[kReduceBlock / 64];
    double v = 0.0, w = 0.0;
    MFV_ROWS(n, i)
    {
        const double ti = t[i];
        v += ti * s[i];
        w += ti * ti;
    }
    const double bv = blockSum(v, lds);
    const double bw = blockSum(w, lds);
    if (threadIdx.x == 0)
    {
        pa[blockIdx.x] = bv;
        pb[blockIdx.x] = bw;
    }
}

__global__ __launch_bounds__(kReduceBlock) void sumMagStopK(int n, const double* __restrict__ a,
                                                            double* __restrict__ partials, const int* __restrict__ fl)
{
    if (fl[BF_STOP]) return;
    __shared__ double lds[kReduceBlock / 64];
    double v = 0.0;
    MFV_ROWS(n, i) v += fabs(a[i]);
    storePartial(blockSum(v, lds), partials);
}

__global__ void dotSeqStopK(int n, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ out,
                            const int* __restrict__ fl)
{
    if (threadIdx.x || blockIdx.x || fl[BF_STOP]) return;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    *out = s;
}

__global__ void sumMagSeqStopK(int n, const double* __restrict__ a, double* __restrict__ out,
                               const int* __restrict__ fl)
{
    if (threadIdx.x || blockIdx.x || fl[BF_STOP]) return;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += fabs(a[i]);
    *out = s;
}

// ---- the recurrence's state steps (pbicgstab.cu:101-145), one thread ----

// after the normFactor and sum|r|: the initial residual; iterate when minIter > 0 or not converged
__global__ void bicgInitK(double* st, int* fl)
{
    if (threadIdx.x || blockIdx.x) return;
    const double init = st[B_T2] / st[B_NF];
    st[B_INIT] = init;
    st[B_FINAL] = init;
    const bool conv = bicgConverged(st, init);
    fl[BF_CONV] = conv ? 1 : 0;
    fl[BF_STOP] = (fl[BF_MINIT] > 0 || !conv) ? 0 : 1;
}

// rA0rA: the breakdown guards and beta
__global__ void bicgRhoK(double* st, int* fl)
{
    if (threadIdx.x || blockIdx.x || fl[BF_STOP]) return;
    const double rhoOld = st[B_RHO];
    const double rho = st[B_T0];
    st[B_RHOOLD] = rhoOld;
    st[B_RHO] = rho;
    if (fabs(rho) < 1e-300)
    {
        fl[BF_STOP] = 1;
        return;
    }
    if (fl[BF_ITER] > 0)
    {
        if (fabs(st[B_OMEGA]) < 1e-300)
        {
            fl[BF_STOP] = 1;
            return;
        }
        st[B_BETA] = (rho / rhoOld) * (st[B_ALPHA] / st[B_OMEGA]);
    }
}

// pA = rA on the first iteration, else rA + beta*(pA - omega*AyA)
__global__ __launch_bounds__(kTpb) void bicgDirK(int n, const double* __restrict__ r, const double* __restrict__ Ay,
                                                 double* __restrict__ p, const double* __restrict__ st,
                                                 const int* __restrict__ fl)
{
    if (fl[BF_STOP]) return;
    const int c = blockIdx.x * kTpb + threadIdx.x;
    if (c >= n) return;
    if (fl[BF_ITER] == 0)
        p[c] = r[c];
    else
        p[c] = r[c] + st[B_BETA] * (p[c] - st[B_OMEGA] * Ay[c]);
}

__global__ void bicgAlphaK(double* st, const int* fl)
{
    if (threadIdx.x || blockIdx.x || fl[BF_STOP]) return;
    st[B_ALPHA] = st[B_RHO] / st[B_T1];
}

__global__ __launch_bounds__(kTpb) void bicgSK(int n, const double* __restrict__ r, const double* __restrict__ Ay,
                                               double* __restrict__ s, const double* __restrict__ st,
                                               const int* __restrict__ fl)
{
    if (fl[BF_STOP]) return;
    const int c = blockIdx.x * kTpb + threadIdx.x;
    if (c < n) s[c] = r[c] - st[B_ALPHA] * Ay[c];
}

// sum|sA|: converged here (past minIter) ends the solve after psi += alpha*yA
__global__ void bicgSResK(double* st, int* fl)
{
    if (threadIdx.x || blockIdx.x || fl[BF_STOP]) return;
    const double fr = st[B_T2] / st[B_NF];
    st[B_FINAL] = fr;
    if (fl[BF_ITER] >= fl[BF_MINIT] && bicgConverged(st, fr))
    {
        fl[BF_CONV] = 1;
        fl[BF_EXIT] = 1;
        fl[BF_STOP] = 1;
    }
}

// the early exit's update, psi += alpha*yA, and its count
__global__ __launch_bounds__(kTpb) void bicgExitK(int n, const double* __restrict__ y, double* __restrict__ psi,
                                                  const double* __restrict__ st, const int* __restrict__ fl)
{
    if (!fl[BF_EXIT]) return;
    const int c = blockIdx.x * kTpb + threadIdx.x;
    if (c < n) psi[c] += st[B_ALPHA] * y[c];
}

__global__ void bicgExitCountK(int* fl)
{
    if (threadIdx.x || blockIdx.x || !fl[BF_EXIT]) return;
    fl[BF_EXIT] = 0;
    fl[BF_ITER] += 1;
}

__global__ void bicgOmegaK(double* st, const int* fl)
{
    if (threadIdx.x || blockIdx.x || fl[BF_STOP]) return;
    st[B_OMEGA] = st[B_T3] / st[B_T4];
}

__global__ __launch_bounds__(kTpb) void bicgUpdK(int n, const double* __restrict__ y, const double* __restrict__ z,
                                                 const double* __restrict__ s, const double* __restrict__ t,
                                                 double* __restrict__ psi, double* __restrict__ r,
                                                 const double* __restrict__ st, const int* __restrict__ fl)
{
    if (fl[BF_STOP]) return;
    const int c = blockIdx.x * kTpb + threadIdx.x;
    if (c >= n) return;
    const double alpha = st[B_ALPHA], omega = st[B_OMEGA];
    psi[c] += alpha * y[c] + omega * z[c];
    r[c] = s[c] - omega * t[c];
}

__global__ void bicgResK(double* st, int* fl)
{
    if (threadIdx.x || blockIdx.x || fl[BF_STOP]) return;
    const double fr = st[B_T2] / st[B_NF];
    st[B_FINAL] = fr;
    const int it = fl[BF_ITER] + 1;
    fl[BF_ITER] = it;
    const bool conv = bicgConverged(st, fr);
    fl[BF_CONV] = conv ? 1 : 0;
    fl[BF_STOP] = ((it < fl[BF_MAXIT] && !conv) || it < fl[BF_MINIT]) ? 0 : 1;
}

struct BicgRun
{
    Ctx& c;
    Mesh& m;
    const BicgPtrs& w;
    LduView A;   // the folded matrix: diag is w.diagC
    int n;
    int gR;

    // out = a.b (or the two sums t.s, t.t into out, out2), stop-aware
    void dot(const double* a, const double* b, double* out, int slot)
    {
        if (c.sequentialReduce)
        {
            dotSeqStopK<<<1, 64, 0, c.stream>>>(n, a, b, out, w.fl);
            return;
        }
        double* pa = c.partials.p + (size_t)slot * kMaxReduceBlocks;
        dotStopK<<<gR, kReduceBlock, 0, c.stream>>>(n, a, b, pa, w.fl);
        finishK<<<1, kFinishBlock, 0, c.stream>>>(gR, pa, StoreTo{out}, w.fl + BF_STOP);
    }
    void dot2(const double* t, const double* s, double* ts, double* tt)
    {
        if (c.sequentialReduce)
        {
            dotSeqStopK<<<1, 64, 0, c.stream>>>(n, t, s, ts, w.fl);
            dotSeqStopK<<<1, 64, 0, c.stream>>>(n, t, t, tt, w.fl);
            return;
        }
        double* pa = c.partials.p + (size_t)3 * kMaxReduceBlocks;
        double* pb = c.partials.p + (size_t)4 * kMaxReduceBlocks;
        dot2StopK<<<gR, kReduceBlock, 0, c.stream>>>(n, t, s, pa, pb, w.fl);
        finishK<<<1, kFinishBlock, 0, c.stream>>>(gR, pa, StoreTo{ts}, w.fl + BF_STOP);
        finishK<<<1, kFinishBlock, 0, c.stream>>>(gR, pb, StoreTo{tt}, w.fl + BF_STOP);
    }
    void sumMag(const double* a, double* out, int slot)
    {
        if (c.sequentialReduce)
        {
            sumMagSeqStopK<<<1, 64, 0, c.stream>>>(n, a, out, w.fl);
            return;
        }
        double* pa = c.partials.p + (size_t)slot * kMaxReduceBlocks;
        sumMagStopK<<<gR, kReduceBlock, 0, c.stream>>>(n, a, pa, w.fl);
        finishK<<<1, kFinishBlock, 0, c.stream>>>(gR, pa, StoreTo{out}, w.fl + BF_STOP);
    }
    void amul(const double* x, double* y)
    {
        const SlabPart part = gatherPart(m);
        if (part.plane)
            bicgAmulK<true><<<gatherBlocks(n, part.plane), kTpb, 0, c.stream>>>(
                n, part, m.nFaces, m.ownerStart.p, m.losortStart.p, m.losort.p, m.losortOwner.p, m.neighbour.p,
                A.diag, A.upper, A.lower, x, y, w.fl);
        else
            bicgAmulK<false><<<blocks(n), kTpb, 0, c.stream>>>(n, part, m.nFaces, m.ownerStart.p, m.losortStart.p,
                                                               m.losort.p, m.losortOwner.p, m.neighbour.p, A.diag,
                                                               A.upper, A.lower, x, y, w.fl);
    }
    void precondition(BicgPrecond pre, const double* r, double* y)
    {
        if (pre == BicgPrecond::Diagonal)
        {
            diagApplyK<<<blocks(n), kTpb, 0, c.stream>>>(n, w.rD, r, y, w.fl);
            return;
        }
        const LevelSched& s = *m.sched;
        for (size_t l = 0; l + 1 < s.fwdStart.size(); ++l)
        {
            const int a = s.fwdStart[l], cnt = s.fwdStart[l + 1] - a;
            diluFwdK<<<blocks(cnt), kTpb, 0, c.stream>>>(cnt, s.fwdCells.p + a, m.losortStart.p, m.losort.p,
                                                         m.losortOwner.p, w.rD, A.lower, r, y, w.fl);
        }
        for (size_t l = 0; l + 1 < s.bwdStart.size(); ++l)
        {
            const int a = s.bwdStart[l], cnt = s.bwdStart[l + 1] - a;
            diluBwdK<<<blocks(cnt), kTpb, 0, c.stream>>>(cnt, s.bwdCells.p + a, m.ownerStart.p, m.neighbour.p,
                                                         w.rD, A.upper, y, w.fl);
        }
    }
    void factor(BicgPrecond pre)
    {
        if (pre == BicgPrecond::Diagonal)
        {
            diagInvK<<<blocks(n), kTpb, 0, c.stream>>>(n, w.diagC, w.rD);
            return;
        }
        const LevelSched& s = *m.sched;
        for (size_t l = 0; l + 1 < s.fwdStart.size(); ++l)
        {
            const int a = s.fwdStart[l], cnt = s.fwdStart[l + 1] - a;
            diluFactorK<<<blocks(cnt), kTpb, 0, c.stream>>>(cnt, s.fwdCells.p + a, m.losortStart.p, m.losort.p,
                                                            m.losortOwner.p, w.diagC, A.upper, A.lower, w.rD);
        }
        invertK<<<blocks(n), kTpb, 0, c.stream>>>(n, w.rD);
    }
};

} // namespace

void BicgWork::size(int n)
{
    if (r.n == (size_t)n) return;
    for (DBuf<double>* v : {&diagC, &b, &rD, &r, &r0, &p, &y, &Ay, &s, &z, &t}) v->resize(n);
    st.resize(B_COUNT);
    fl.resize(BF_COUNT);
}

namespace {

// The solve itself, in the workspace w (every vector n long; st and fl the state). Nothing here allocates:
// the level schedule (DILU) must already be on the mesh.
PcgPerf bicgRunSolve(Ctx& c, Mesh& m, const LduView& A0, const double* source, const double* iC, const double* bC,
                     double* psi, const BicgControls& bc, const BicgPtrs& w)
{
    const int n = m.nCells;
    BicgRun R{c, m, w, LduView{n, m.nFaces, w.diagC, A0.upper, A0.lower}, n, reduceGrid(n)};
    double* S = w.st;
    int* F = w.fl;
    bicgSetupK<<<1, 1, 0, c.stream>>>(S, F, bc.tolerance, bc.relTol, bc.maxIter, bc.minIter);
    bicgFoldK<<<blocks(n), kTpb, 0, c.stream>>>(n, m.bndStart.p, m.bndFaces.p, A0.diag, source, iC, bC, w.diagC,
                                                w.b);
    R.factor(bc.precond);
    // rA = b - A psi (Amul, then the difference), rA0 = rA; normFactor with the folded diagonal
    R.amul(psi, w.y);
    bicgSubK<<<blocks(n), kTpb, 0, c.stream>>>(n, w.b, w.y, w.r, w.r0);
    launchNormFactor(c, m, R.A, psi, w.b, S + B_NF);
    R.sumMag(w.r, S + B_T2, 2);
    bicgInitK<<<1, 1, 0, c.stream>>>(S, F);
    hipCheck(hipGetLastError(), "pbicgstab init");
    // the reference's do-while runs its body at least once when not converged on entry
    const int limit = std::max(1, std::max(bc.maxIter, bc.minIter));
    for (int it = 0; it < limit; ++it)
    {
        R.dot(w.r0, w.r, S + B_T0, 0);
        bicgRhoK<<<1, 1, 0, c.stream>>>(S, F);
        bicgDirK<<<blocks(n), kTpb, 0, c.stream>>>(n, w.r, w.Ay, w.p, S, F);
        R.precondition(bc.precond, w.p, w.y);
        R.amul(w.y, w.Ay);
        R.dot(w.r0, w.Ay, S + B_T1, 1);
        bicgAlphaK<<<1, 1, 0, c.stream>>>(S, F);
        bicgSK<<<blocks(n), kTpb, 0, c.stream>>>(n, w.r, w.Ay, w.s, S, F);
        R.sumMag(w.s, S + B_T2, 2);
        bicgSResK<<<1, 1, 0, c.stream>>>(S, F);
        bicgExitK<<<blocks(n), kTpb, 0, c.stream>>>(n, w.y, psi, S, F);
        bicgExitCountK<<<1, 1, 0, c.stream>>>(F);
        R.precondition(bc.precond, w.s, w.z);
        R.amul(w.z, w.t);
        R.dot2(w.t, w.s, S + B_T3, S + B_T4);
        bicgOmegaK<<<1, 1, 0, c.stream>>>(S, F);
        bicgUpdK<<<blocks(n), kTpb, 0, c.stream>>>(n, w.y, w.z, w.s, w.t, psi, w.r, S, F);
        R.sumMag(w.r, S + B_T2, 2);
        bicgResK<<<1, 1, 0, c.stream>>>(S, F);
        hipCheck(hipGetLastError(), "pbicgstab iteration");
        if ((it + 1) % std::max(1, bc.checkEvery) == 0 || it + 1 == limit)
        {
            hipCheck(hipMemcpyAsync(c.mailbox, F, sizeof(int) * BF_COUNT, hipMemcpyDeviceToHost, c.stream), "poll");
            hipCheck(hipStreamSynchronize(c.stream), "poll sync");
            if (reinterpret_cast<const int*>(c.mailbox)[BF_STOP]) break;
        }
    }
    hipCheck(hipMemcpyAsync(c.mailbox, S, sizeof(double) * B_COUNT, hipMemcpyDeviceToHost, c.stream), "report");
    hipCheck(hipMemcpyAsync(c.mailbox + B_COUNT, F, sizeof(int) * BF_COUNT, hipMemcpyDeviceToHost, c.stream),
             "report flags");
    hipCheck(hipStreamSynchronize(c.stream), "report sync");
    const double* st = c.mailbox;
    const int* fl = reinterpret_cast<const int*>(c.mailbox + B_COUNT);
    PcgPerf perf;
    perf.initialResidual = st[B_INIT];
    perf.finalResidual = st[B_FINAL];
    perf.nIterations = fl[BF_ITER];
    perf.converged = fl[BF_CONV] != 0;
    return perf;
}


} // namespace

PcgPerf pbicgstabSolve(Ctx& c, Mesh& m, const LduView& A0, const double* source, const double* iC, const double* bC,
                       double* psi, const BicgControls& bc)
{
    if (bc.maxIter < 0 || bc.minIter < 0) invalid("pbicgstab: negative iteration limits");
    if (m.nBFaces && (!iC || !bC)) invalid("pbicgstab: boundary coefficients missing");
    const int n = m.nCells;
    if (!m.bicg) m.bicg.reset(new BicgWork());
    BicgWork& w = *m.bicg;
    w.size(n);
    if (bc.precond == BicgPrecond::DILU) buildLevelSchedule(c, m);
    const BicgPtrs ptrs{w.diagC.p, w.b.p, w.rD.p, w.r.p, w.r0.p, w.p.p, w.y.p, w.Ay.p, w.s.p, w.z.p, w.t.p,
                        w.st.p, w.fl.p};
    return bicgRunSolve(c, m, A0, source, iC, bC, psi, bc, ptrs);
}

// ---- the no-allocation variant: the workspace in one static device array ----
//
// scratch holds the eleven work vectors of a solve of up to kStaticBicgMaxCells cells, each at a 256-byte
// aligned offset, then the solver state (B_COUNT doubles) and its flags (BF_COUNT ints). Each GPU the
// library runs on holds its own copy until the process ends. In host code `scratch` names a host-side
// stand-in, so its device address on each GPU is looked up once, with that GPU current, and kept in
// gStaticScratch; only that address is handed to kernels.
namespace {
constexpr size_t kStaticVec = kStaticBicgMaxCells;      // doubles per vector: a multiple of 32 (256 bytes)
constexpr size_t kStaticState = kStaticVec * 11;        // offset of the state, then the flags 256 bytes on
constexpr size_t kStaticDoubles = kStaticState + 64;
static_assert(kStaticVec % 32 == 0, "vectors stay 256-byte aligned");
static_assert(B_COUNT <= 32 && BF_COUNT * sizeof(int) <= 32 * sizeof(double), "state fits its slot");
constexpr int kMaxDevices = 64;
std::mutex gStaticMutex;
double* gStaticScratch[kMaxDevices] = {};
} // namespace

static __device__ __align__(256) double scratch[kStaticDoubles];

void staticBicgScratchInit(int device)
{
    if (device < 0 || device >= kMaxDevices) invalid("static workspace: device ordinal out of range");
    std::lock_guard<std::mutex> lk(gStaticMutex);
    if (gStaticScratch[device]) return;
    void* ptr = nullptr;
    hipCheck(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "static workspace address");
    if (!ptr) invalid("static workspace: no device address");
    gStaticScratch[device] = static_cast<double*>(ptr);
}

PcgPerf pbicgstabSolveStatic(Ctx& c, Mesh& m, const LduView& A0, const double* source, const double* iC,
                             const double* bC, double* psi, const BicgControls& bc)
{
    if (bc.maxIter < 0 || bc.minIter < 0) invalid("pbicgstab (no allocation): negative iteration limits");
    if (m.nBFaces && (!iC || !bC)) invalid("pbicgstab (no allocation): boundary coefficients missing");
    const int n = m.nCells;
    if ((size_t)n > kStaticVec)
        invalid("pbicgstab (no allocation): " + std::to_string(n) + " cells exceed the static workspace's " +
                std::to_string(kStaticBicgMaxCells));
    if (bc.precond == BicgPrecond::DILU && !m.sched)
        invalid("pbicgstab (no allocation): DILU needs the mesh's level schedule; a first mfv_pbicgstab_solve, "
                "mfv_pcg_dic_solve or sweep on this mesh builds it (or use the diagonal preconditioner)");
    double* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(gStaticMutex);
        if (c.device >= 0 && c.device < kMaxDevices) base = gStaticScratch[c.device];
    }
    if (!base) invalid("pbicgstab (no allocation): the static workspace of this device was not set up");
    double* v[11];
    for (int k = 0; k < 11; ++k) v[k] = base + (size_t)k * kStaticVec;
    double* st = base + kStaticState;
    int* fl = reinterpret_cast<int*>(base + kStaticState + 32);
    const BicgPtrs ptrs{v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8], v[9], v[10], st, fl};
    return bicgRunSolve(c, m, A0, source, iC, bC, psi, bc, ptrs);
}

} // namespace mfv
