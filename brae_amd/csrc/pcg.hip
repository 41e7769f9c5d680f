// AMG-preconditioned conjugate gradients, device-resident.
//
// Reference: brae's deviceAMGPCG (src/matrices/lduMatrix/solvers/PCG/device_amg_pcg.cu:183-411,
// 638-848): OpenFOAM's residual norm (sum|r| / normFactor), Fletcher-Reeves beta = wArA/wArAold
// with the 1e-300 guard (pcgSearchDirK :60-77), or the flexible Polak-Ribiere+ beta
// max(0, (z.r - z.r_old)/rho_old) (flexBetaK :37-49) for a preconditioner that is not a fixed linear
// operator (a reduced-precision cycle); stop when res < tol, or res < relTol*init, or at maxIter, never
// before minIter (pcgSetCondK :164-180).
//
// HIP has no conditional graph nodes, so the loop is host-driven but never waits on the device per
// iteration: the stop decision is taken ON THE DEVICE by the last workgroup of the update reduction,
// and every kernel of an iteration returns at once when the stop flag is set. The host enqueues
// iterations back to back and reads the flag through a pinned mailbox every `checkEvery` iterations;
// iterations enqueued past convergence do no work, so the result does not depend on checkEvery.
#include "amg.hpp"
#include "reduce.hpp"
#include "sell_device.hpp"

namespace mfv {

namespace {

constexpr int kTpb = 256;
inline int blocks(long long n) { return (int)((n + kTpb - 1) / kTpb); }

// device state (doubles)
enum : int { S_RHO = 0, S_PQ, S_RES, S_INIT, S_NF, S_BETA, S_SUMPSI, S_FINAL, S_TOL, S_RELTOL, S_COUNT = 16 };
// device flags (ints)
enum : int { F_ITER = 0, F_STOP, F_CONV, F_MAXIT, F_MINIT, F_COUNT = 8 };

struct PcgDev
{
    double* st;
    int* fl;
};

__device__ __forceinline__ bool convergedAt(const double* st, double res)
{
    return res < st[S_TOL] || (st[S_RELTOL] > 0.0 && res < st[S_RELTOL] * st[S_INIT]);
}

__global__ __launch_bounds__(kReduceBlock) void sumPsiK(int n, const double* __restrict__ a, double* partials)
{
    __shared__ double lds[kReduceBlock / 64];
    double v = 0.0;
    MFV_ROWS(n, i) v += a[i];
    storePartial(blockSum(v, lds), partials);
}

__device__ __forceinline__ void storePartials2(double a, double b, double* pA, double* pB)
{
    if (threadIdx.x == 0)
    {
        pA[blockIdx.x] = a;
        pB[blockIdx.x] = b;
    }
}

// r = b - A psi; sum|r| and (without a caller normFactor) the normFactor summand
// |A psi - sumA*xRef| + |b - sumA*xRef|, sumA = A.1 in face order (device_pcg.cu:118-170).
struct InitPost
{
    PcgDev d;
    int haveNf;
    const double* nfIn;
    __device__ void operator()(double sumR, double sumNf) const
    {
        const double nf = haveNf ? *nfIn : sumNf + 1e-20;
        d.st[S_NF] = nf;
        const double init = sumR / nf;
        d.st[S_INIT] = init;
        d.st[S_FINAL] = init;
        d.st[S_RHO] = 1e300;
        d.fl[F_ITER] = 0;
        const bool conv = convergedAt(d.st, init);
        d.fl[F_CONV] = conv ? 1 : 0;
        d.fl[F_STOP] = (d.fl[F_MINIT] <= 0 && conv) ? 1 : 0;
    }
};

__global__ __launch_bounds__(kReduceBlock) void pcgInitK(SellView S, const double* __restrict__ val,
                                                         const double* __restrict__ diag,
                                                         const double* __restrict__ psi, const double* __restrict__ b,
                                                         double* __restrict__ r, PcgDev d, int haveNf,
                                                         double* pA, double* pB, double* seqA, double* seqB)
{
    __shared__ double lds[kReduceBlock / 64];
    const double xRef = haveNf ? 0.0 : d.st[S_SUMPSI] / (double)S.n;
    double sr = 0.0, snf = 0.0;
    MFV_ROWS(S.n, c)
    {
        const double ax = rowOffSum<double, double>(S, c, val, PlainX<double>{psi}, diag[c] * psi[c]);
        const double rc = b[c] - ax;
        r[c] = rc;
        double e = 0.0;
        if (!haveNf)
        {
            const double sa = rowOffSum<double, double>(S, c, val, [](int) { return 1.0; }, diag[c]);
            const double t = sa * xRef;
            e = fabs(ax - t) + fabs(b[c] - t);
        }
        if (seqA)
        {
            seqA[c] = fabs(rc);
            seqB[c] = e;
        }
        else
        {
            sr += fabs(rc);
            snf += e;
        }
    }
    if (seqA) return;
    const double a = blockSum(sr, lds);
    const double bsum = blockSum(snf, lds);
    storePartials2(a, bsum, pA, pB);
}

__global__ void seqSumK(int n, const double* __restrict__ a, double* __restrict__ out)
{
    if (threadIdx.x || blockIdx.x) return;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i];
    *out = s;
}

// rho = z.r (and z.r_old for the flexible beta); the last workgroup forms beta
struct ZrPost
{
    PcgDev d;
    int flexible;
    __device__ void operator()(double zr, double zrOld) const
    {
        const double prev = d.st[S_RHO];
        double beta = 0.0;
        if (d.fl[F_ITER] > 0)
        {
            if (flexible)
                beta = (prev > 1e-300 || prev < -1e-300) ? fmax(0.0, (zr - zrOld) / prev) : 0.0;
            else
                beta = fabs(prev) > 1e-300 ? zr / prev : 0.0;
        }
        d.st[S_BETA] = beta;
        d.st[S_RHO] = zr;
    }
};

// the V-cycle's result z at row i in FP64: the stored value, times the FP16 cycle's input scale
template <typename Tv, typename Ts>
struct CycleOut
{
    CycVec<Tv, Ts> z;
    double sig;
    __device__ double operator()(int i) const { return sig != 0.0 ? (double)z(i) * sig : (double)z(i); }
};

template <typename Tv, typename Ts>
__device__ __forceinline__ CycleOut<Tv, Ts> cycleOut(const Ts* z, const float* xm, const double* sig)
{
    return CycleOut<Tv, Ts>{CycVec<Tv, Ts>(z, xm), sig ? *sig : 0.0};
}

// Flex (the flexible beta's z.r_old) is a template parameter: a load behind a run-time flag is issued
// behind a branch, apart from the row's other loads.
template <typename Tv, typename Ts, bool Flex>
__global__ __launch_bounds__(kReduceBlock) void pcgZrK(int n, const Ts* __restrict__ zs, const float* __restrict__ xm,
                                                       const double* __restrict__ sig, const double* __restrict__ r,
                                                       const double* __restrict__ rOld, PcgDev d, double* pA,
                                                       double* pB, double* seqA, double* seqB)
{
    if (d.fl[F_STOP]) return;
    __shared__ double lds[kReduceBlock / 64];
    const CycleOut<Tv, Ts> z = cycleOut<Tv, Ts>(zs, xm, sig);
    double a = 0.0, o = 0.0;
    MFV_ROWS(n, i)
    {
        const double zi = z(i);
        const double ea = zi * r[i];
        const double eo = Flex ? zi * rOld[i] : 0.0;
        if (seqA)
        {
            seqA[i] = ea;
            seqB[i] = eo;
        }
        else
        {
            a += ea;
            o += eo;
        }
    }
    if (seqA) return;
    const double sa = blockSum(a, lds);
    const double so = blockSum(o, lds);
    storePartials2(sa, so, pA, pB);
}

template <typename Tv, typename Ts>
struct DirX
{
    CycleOut<Tv, Ts> z;
    const double* __restrict__ pOld;
    double beta;
    int first;
    // pOld is loaded on the first iteration too (its buffer is allocated; the value is discarded), so the
    // loads of a row's neighbours issue as one batch instead of behind a branch each
    __device__ double operator()(int j) const
    {
        const double po = pOld[j];
        const double zj = z(j);
        const double pj = beta * po + zj;
        return first ? zj : pj;
    }
};

// p = z (first iteration) or beta*p + z; q = A p; p.q
struct PqPost
{
    PcgDev d;
    __device__ void operator()(double s) const { d.st[S_PQ] = s; }
};

template <typename Tv, typename Ts>
__global__ __launch_bounds__(kReduceBlock) void pcgDirAmulK(SellView S, const double* __restrict__ val,
                                                            const double* __restrict__ diag, const Ts* __restrict__ z,
                                                            const float* __restrict__ xm,
                                                            const double* __restrict__ sig,
                                                            const double* __restrict__ pOld,
                                                            double* __restrict__ pNew, double* __restrict__ q, PcgDev d,
                                                            double* partials, double* seqA)
{
    if (d.fl[F_STOP]) return;
    __shared__ double lds[kReduceBlock / 64];
    const DirX<Tv, Ts> P{cycleOut<Tv, Ts>(z, xm, sig), pOld, d.st[S_BETA], d.fl[F_ITER] == 0 ? 1 : 0};
    double s = 0.0;
    MFV_ROWS(S.n, c)
    {
        const double pc = P(c);
        const double ax = rowOffSum<double, double>(S, c, val, P, ldStream(diag + c) * pc);
        __builtin_nontemporal_store(pc, pNew + c);
        __builtin_nontemporal_store(ax, q + c);
        if (seqA)
            seqA[c] = ax * pc;
        else
            s += ax * pc;
    }
    if (seqA) return;
    storePartial(blockSum(s, lds), partials);
}

// Grid 0 as diagonals (a fixed stencil of M positive offsets), by value.
struct DiaView
{
    int n, pad;
    int off[kDiaMax];
};

// The same product on grid 0's diagonal copy, each thread on R rows at once (c, c + kReduceBlock, ...)
// with every load of all R issued before the sums. Row c sums D p_c, then its lower entries at columns
// c - off[M-1] .. c - off[0], then its upper ones at c + off[0] .. c + off[M-1]: its face order, with a
// zero coefficient (times a finite p) where a face is absent. The upper coefficients stream once
// (nontemporal); the lower ones are the rows above's upper coefficients, reread through the cache. A
// column outside the grid reads the row itself.
template <int R, int M, typename XF>
__device__ __forceinline__ void diaRows(const DiaView& D, int c0, int end, const double* __restrict__ up,
                                        const double* __restrict__ diag, XF P, double* ax, double* pc)
{
    double a[R][2 * M], xv[R][2 * M], dg[R];
#pragma unroll
    for (int i = 0; i < R; ++i)
    {
        const int ci = c0 + i * kReduceBlock;
        const int c = ci < end ? ci : c0;
        dg[i] = ldStream(diag + c);
        int col[2 * M];
#pragma unroll
        for (int k = 0; k < M; ++k)
        {
            const int kk = M - 1 - k;
            const int j = c - D.off[kk];
            const bool in = j >= 0;
            col[k] = in ? j : c;
            const double v = up[(size_t)kk * D.pad + col[k]];
            a[i][k] = in ? v : 0.0;
        }
#pragma unroll
        for (int k = 0; k < M; ++k)
        {
            const int j = c + D.off[k];
            col[M + k] = j < D.n ? j : c;
            a[i][M + k] = ldStream(up + (size_t)k * D.pad + c);
        }
        pc[i] = P(c);
#pragma unroll
        for (int k = 0; k < 2 * M; ++k) xv[i][k] = P(col[k]);
    }
#pragma unroll
    for (int i = 0; i < R; ++i)
    {
        double acc = dg[i] * pc[i];
#pragma unroll
        for (int k = 0; k < 2 * M; ++k) acc += a[i][k] * xv[i][k];
        ax[i] = acc;
    }
}

// One row per thread: with a slab partition (part.plane > 0, launched on slabBlocks) its rows, else
// the contiguous XCD-ordered chunks of a reduceGridRows(n, 1) launch.
template <int M, typename Tv, typename Ts>
__global__ __launch_bounds__(kReduceBlock) void pcgDirDiaK(DiaView D, SlabPart part, const double* __restrict__ up,
                                                           const double* __restrict__ diag, const Ts* __restrict__ z,
                                                           const float* __restrict__ xm,
                                                           const double* __restrict__ sig,
                                                           const double* __restrict__ pOld,
                                                           double* __restrict__ pNew, double* __restrict__ q, PcgDev d,
                                                           double* partials, double* seqA)
{
    if (d.fl[F_STOP]) return;
    __shared__ double lds[kReduceBlock / 64];
    const DirX<Tv, Ts> P{cycleOut<Tv, Ts>(z, xm, sig), pOld, d.st[S_BETA], d.fl[F_ITER] == 0 ? 1 : 0};
    double s = 0.0;
    auto row = [&](int c) {
        double ax[1], pc[1];
        diaRows<1, M>(D, c, c + 1, up, diag, P, ax, pc);
        __builtin_nontemporal_store(pc[0], pNew + c);
        __builtin_nontemporal_store(ax[0], q + c);
        if (seqA)
            seqA[c] = ax[0] * pc[0];
        else
            s += ax[0] * pc[0];
    };
    if (part.plane)
    {
        for (int r = 0; r < part.rpt; ++r)
        {
            const int c = slabRow(part, r);
            if (c >= 0) row(c);
        }
    }
    else
    {
        MFV_ROWS(D.n, c) row(c);
    }
    if (seqA) return;
    storePartial(blockSum(s, lds), partials);
}

// psi += alpha p, r -= alpha q, sum|r|; the stop decision; the next cycle's grid-0 entry. The FP16
// cycle's input scale rolls on: the update stored the next cycle's input at sig[1], which becomes that
// cycle's scale sig[0]; sig[1] becomes the power of two at or above the new max|r|.
struct UpdPost
{
    PcgDev d;
    double* sig;
    __device__ void roll(double maxR) const
    {
        if (!sig) return;
        sig[0] = sig[1];
        if (maxR > 0.0) sig[1] = exp2(ceil(log2(maxR)));
    }
    __device__ void operator()(double sumR) const
    {
        const double res = sumR / d.st[S_NF];
        d.st[S_FINAL] = res;
        const int it = d.fl[F_ITER] + 1;
        d.fl[F_ITER] = it;
        const bool conv = convergedAt(d.st, res);
        d.fl[F_CONV] = conv ? 1 : 0;
        d.fl[F_STOP] = ((conv || it >= d.fl[F_MAXIT]) && it >= d.fl[F_MINIT]) ? 1 : 0;
    }
};

// with `sig` (the FP16 cycle) the next cycle's input is r/sig[1], and max|r| goes to pMax. The residual
// goes from rIn to rOut: the two buffers alternate, so rIn is the flexible beta's r_old one iteration on
// (no copy of r).
template <typename Tv, typename Ts>
__global__ __launch_bounds__(kReduceBlock) void pcgUpdateK(int n, const double* __restrict__ p,
                                                           const double* __restrict__ q, double* __restrict__ psi,
                                                           const double* __restrict__ rIn, double* __restrict__ rOut,
                                                           const Tv* __restrict__ diag0,
                                                           Ts* __restrict__ b0, Ts* __restrict__ x0,
                                                           const float* __restrict__ xm, const double* __restrict__ sig,
                                                           int fromZero, PcgDev d, double* partials, double* pMax,
                                                           double* seqA)
{
    if (d.fl[F_STOP]) return;
    __shared__ double lds[kReduceBlock / 64];
    const double pq = d.st[S_PQ];
    const double alpha = fabs(pq) > 1e-300 ? d.st[S_RHO] / pq : 0.0;
    const double isig = sig ? 1.0 / sig[1] : 1.0;
    const CycVec<Tv, Ts> B(b0, nullptr), X(x0, xm);
    double s = 0.0, mr = 0.0;
    MFV_ROWS(n, i)
    {
        psi[i] += alpha * p[i];
        const double rn = rIn[i] - alpha * q[i];
        rOut[i] = rn;
        if (seqA)
            seqA[i] = fabs(rn);
        else
            s += fabs(rn);
        mr = fmax(mr, fabs(rn));
        const Tv bv = sig ? (Tv)(rn * isig) : (Tv)rn;
        B.set(i, bv);
        X.set(i, fromZero ? Tv(0) + Tv(kOmega) * (bv - Tv(0)) / safeDiag(diag0[i]) : Tv(0));
    }
    if (pMax)
    {
        const double m = blockMax(mr, lds);
        if (threadIdx.x == 0) pMax[blockIdx.x] = m;
    }
    if (seqA) return;
    storePartial(blockSum(s, lds), partials);
}

// Level one of a two-level finish, for the outer products' tens of thousands of partials: workgroup b sums
// (and with Max takes the maximum of) the b-th of gridDim.x contiguous ranges of them, in a fixed order, so
// the one-workgroup finish then reads gridDim.x values instead of reading them all on one CU.
constexpr int kLevelOneBlocks = 64;
constexpr int kLevelOneFrom = 8192;   // partials at and above which the finish takes two levels
template <bool Max>
__global__ __launch_bounds__(kFinishBlock) void finishLevelOneK(int nParts, const double* __restrict__ pa,
                                                                const double* __restrict__ pm,
                                                                double* __restrict__ outA, double* __restrict__ outM,
                                                                const int* __restrict__ stop)
{
    if (stop && *stop) return;
    __shared__ double lds[kFinishBlock / 64];
    const int per = (nParts + (int)gridDim.x - 1) / (int)gridDim.x;
    const int lo = (int)blockIdx.x * per, hi = min(nParts, lo + per);
    double sa = 0.0, m = 0.0;
    for (int i = lo + (int)threadIdx.x; i < hi; i += kFinishBlock)
    {
        sa += pa[i];
        if (Max) m = fmax(m, pm[i]);
    }
    sa = blockSum<kFinishBlock>(sa, lds);
    if (Max) m = blockMax<kFinishBlock>(m, lds);
    if (threadIdx.x == 0)
    {
        outA[blockIdx.x] = sa;
        if (Max) outM[blockIdx.x] = m;
    }
}

// the update's finish with the FP16 cycle's max|r|
__global__ __launch_bounds__(kFinishBlock) void updFinishK(int nParts, const double* __restrict__ pa,
                                                           const double* __restrict__ pMax, UpdPost post,
                                                           const int* __restrict__ stop)
{
    if (stop && *stop) return;
    __shared__ double lds[kFinishBlock / 64];
    double h, mp;
    finishShareMax(nParts, pa, pMax, h, mp);
    const double s = blockSum<kFinishBlock>(h, lds);
    const double m = blockMax<kFinishBlock>(mp, lds);
    if (threadIdx.x == 0)
    {
        post(s);
        post.roll(m);
    }
}

// the sequential-order finish of the update with the FP16 cycle's max|r| (one thread)
__global__ void updSeqFinishK(int n, const double* __restrict__ a, int nParts, const double* __restrict__ pMax,
                              UpdPost post, const int* __restrict__ stop)
{
    if (threadIdx.x || blockIdx.x) return;
    if (stop && *stop) return;
    double s = 0.0, m = 0.0;
    for (int i = 0; i < n; ++i) s += a[i];
    for (int i = 0; i < nParts; ++i) m = fmax(m, pMax[i]);
    post(s);
    post.roll(m);
}

__global__ void pcgSetupK(PcgDev d, double tol, double relTol, int maxIter, int minIter)
{
    if (threadIdx.x || blockIdx.x) return;
    d.st[S_TOL] = tol;
    d.st[S_RELTOL] = relTol;
    d.fl[F_MAXIT] = maxIter;
    d.fl[F_MINIT] = minIter;
    d.fl[F_STOP] = 0;
    d.fl[F_ITER] = 0;
}

template <typename Tv, typename Ts>
PcgPerf solveT(Ctx& c, const Mesh& m, Amg& a, const double* diag, const double* b, double* psi,
               const double* normFactor, const PcgControls& pc, Ts* b0, Ts* x0, const Tv* diag0)
{
    // the FP16 cycle: grid 0's solution-vector multiplier and the input scales
    const bool half = pc.prec == Prec::F16;
    const float* xm0 = half ? a.xmul.p : nullptr;
    double* sig = half ? a.sig.p : nullptr;
    double* pMax = half ? c.partials.p + 7 * (size_t)kMaxReduceBlocks : nullptr;
    const int n = m.nCells;
    PcgWork& w = *a.work;
    PcgDev d{w.st.p, w.fl.p};
    const Grid& g0 = a.grids[0];
    const SellView S = viewOf(g0.sell);
    const int gR = reduceGrid(n);
    // rows per thread, measured per kernel on the 20 M channel: the FP64 direction product 4 (1 on the
    // diagonal copy; 2 measured 5 % slower), z.r 8, the update 2 (reduceGrid; 4 measured 5 % slower, more than
    // its finish saves)
    const bool dia = g0.diaM > 0;
    DiaView D{n, g0.diaPad, {}};
    for (int k = 0; k < kDiaMax; ++k) D.off[k] = g0.diaOff[k];
    // the diagonal product: one row per thread, on plane slabs when the stencil's planes are wide enough
    const int plane = dia ? g0.diaOff[g0.diaM - 1] : 0;
    const SlabPart part = plane >= kSlabMinPlane ? slabPart(n, plane) : SlabPart{n, 0, 1};
    const int gRd = !dia ? reduceGridRows(n, 4) : part.plane ? slabBlocks(n, part.plane, part.rpt) : reduceGridRows(n, 1);
    const int gRz = reduceGridRows(n, 8);
    const int gRu = reduceGrid(n);
    double* pa = c.partials.p + 5 * (size_t)kMaxReduceBlocks;
    double* pb = c.partials.p + 6 * (size_t)kMaxReduceBlocks;
    // a two-level finish's level-one sums and maxima: the top of pb's region, beyond any partition's
    // partials (pb holds the z.r partials, at most reduceGridRows(n, 8) of them, when they are read)
    double* l1a = c.partials.p + 7 * (size_t)kMaxReduceBlocks - 2 * kLevelOneBlocks;
    double* l1m = l1a + kLevelOneBlocks;
    const bool seq = c.sequentialReduce;
    if (seq && w.seqA.n != (size_t)n)
    {
        w.seqA.resize(n);
        w.seqB.resize(n);
    }
    double* sA = seq ? w.seqA.p : nullptr;
    double* sB = seq ? w.seqB.p : nullptr;
    const int* stopF = w.fl.p + F_STOP;
    pcgSetupK<<<1, 64, 0, c.stream>>>(d, pc.tolerance, pc.relTol, pc.maxIter, pc.minIter);
    if (!normFactor)
    {
        if (seq)
            seqSumK<<<1, 64, 0, c.stream>>>(n, psi, w.st.p + S_SUMPSI);
        else
        {
            sumPsiK<<<gR, kReduceBlock, 0, c.stream>>>(n, psi, pa);
            finishK<<<1, kFinishBlock, 0, c.stream>>>(gR, pa, StoreTo{w.st.p + S_SUMPSI}, nullptr);
        }
    }
    pcgInitK<<<gR, kReduceBlock, 0, c.stream>>>(S, g0.off64.p, diag, psi, b, w.r.p, d, normFactor ? 1 : 0, pa, pb,
                                               sA, sB);
    const InitPost initPost{d, normFactor ? 1 : 0, normFactor};
    if (seq)
        seqFinish2K<<<1, 64, 0, c.stream>>>(n, sA, sB, initPost, nullptr);
    else
        finish2K<<<1, kFinishBlock, 0, c.stream>>>(gR, pa, pb, initPost, nullptr);
    hipCheck(hipGetLastError(), "pcg init");
    VcycleControls vc;
    vc.nPre = pc.nPre;
    vc.nPost = pc.nPost;
    vc.scale = pc.scaleCorrection;
    amgStartFrom(c, a, w.r.p, vc, w.fl.p + F_STOP);
    double* p[2] = {w.p0.p, w.p1.p};
    // the residual of iteration it is in R[it & 1], the previous one (the flexible beta's r_old) in the
    // other, which the update then overwrites with the next residual
    double* R[2] = {w.r.p, w.rOld.p};
    // one iteration, enqueued; `it` selects the direction and residual buffers (they alternate)
    auto iteration = [&](int it) {
        const Ts* z;
        {
            ProfScope ps(c, kProfVcycle);
            z = static_cast<const Ts*>(amgVcycle(c, a, vc, w.fl.p + F_STOP));
        }
        {
            ProfScope ps(c, kProfPcgZr);
            if (pc.flexible)
                pcgZrK<Tv, Ts, true><<<gRz, kReduceBlock, 0, c.stream>>>(n, z, xm0, sig, R[it & 1], R[(it + 1) & 1], d,
                                                                        pa, pb, sA, sB);
            else
                pcgZrK<Tv, Ts, false><<<gRz, kReduceBlock, 0, c.stream>>>(n, z, xm0, sig, R[it & 1], R[(it + 1) & 1],
                                                                         d, pa, pb, sA, sB);
        }
        {
            const ZrPost post{d, pc.flexible ? 1 : 0};
            if (seq)
                seqFinish2K<<<1, 64, 0, c.stream>>>(n, sA, sB, post, stopF);
            else
                finish2K<<<1, kFinishBlock, 0, c.stream>>>(gRz, pa, pb, post, stopF);
        }
        {
            ProfScope ps(c, kProfPcgSpmv64);
            const double* pOld = p[it & 1];
            double* pNew = p[(it + 1) & 1];
            switch (dia ? g0.diaM : 0)
            {
            case 1:
                pcgDirDiaK<1, Tv, Ts><<<gRd, kReduceBlock, 0, c.stream>>>(D, part, g0.dia64.p, diag, z, xm0, sig, pOld, pNew,
                                                                         w.q.p, d, pa, sA);
                break;
            case 2:
                pcgDirDiaK<2, Tv, Ts><<<gRd, kReduceBlock, 0, c.stream>>>(D, part, g0.dia64.p, diag, z, xm0, sig, pOld, pNew,
                                                                         w.q.p, d, pa, sA);
                break;
            case 3:
                pcgDirDiaK<3, Tv, Ts><<<gRd, kReduceBlock, 0, c.stream>>>(D, part, g0.dia64.p, diag, z, xm0, sig, pOld, pNew,
                                                                         w.q.p, d, pa, sA);
                break;
            case 4:
                pcgDirDiaK<4, Tv, Ts><<<gRd, kReduceBlock, 0, c.stream>>>(D, part, g0.dia64.p, diag, z, xm0, sig, pOld, pNew,
                                                                         w.q.p, d, pa, sA);
                break;
            default:
                pcgDirAmulK<Tv, Ts><<<gRd, kReduceBlock, 0, c.stream>>>(S, g0.off64.p, diag, z, xm0, sig, pOld, pNew,
                                                                       w.q.p, d, pa, sA);
            }
        }
        if (seq)
            seqFinishK<<<1, 64, 0, c.stream>>>(n, sA, PqPost{d}, stopF);
        else if (gRd >= kLevelOneFrom)
        {
            finishLevelOneK<false><<<kLevelOneBlocks, kFinishBlock, 0, c.stream>>>(gRd, pa, nullptr, l1a, nullptr,
                                                                                  stopF);
            finishK<<<1, kFinishBlock, 0, c.stream>>>(kLevelOneBlocks, l1a, PqPost{d}, stopF);
        }
        else
            finishK<<<1, kFinishBlock, 0, c.stream>>>(gRd, pa, PqPost{d}, stopF);
        {
            ProfScope ps(c, kProfPcgUpdate);
            pcgUpdateK<Tv, Ts><<<gRu, kReduceBlock, 0, c.stream>>>(n, p[(it + 1) & 1], w.q.p, psi, R[it & 1],
                                                                 R[(it + 1) & 1], diag0, b0, x0, xm0, sig,
                                                                 pc.nPre > 0 ? 1 : 0, d, pa, pMax, sA);
        }
        const UpdPost upd{d, sig};
        if (half && seq)
            updSeqFinishK<<<1, 64, 0, c.stream>>>(n, sA, gRu, pMax, upd, stopF);
        else if (half && gRu >= kLevelOneFrom)
        {
            finishLevelOneK<true><<<kLevelOneBlocks, kFinishBlock, 0, c.stream>>>(gRu, pa, pMax, l1a, l1m, stopF);
            updFinishK<<<1, kFinishBlock, 0, c.stream>>>(kLevelOneBlocks, l1a, l1m, upd, stopF);
        }
        else if (half)
            updFinishK<<<1, kFinishBlock, 0, c.stream>>>(gRu, pa, pMax, upd, stopF);
        else if (seq)
            seqFinishK<<<1, 64, 0, c.stream>>>(n, sA, upd, stopF);
        else if (gRu >= kLevelOneFrom)
        {
            finishLevelOneK<false><<<kLevelOneBlocks, kFinishBlock, 0, c.stream>>>(gRu, pa, nullptr, l1a, nullptr,
                                                                                  stopF);
            finishK<<<1, kFinishBlock, 0, c.stream>>>(kLevelOneBlocks, l1a, upd, stopF);
        }
        else
            finishK<<<1, kFinishBlock, 0, c.stream>>>(gRu, pa, upd, stopF);
        hipCheck(hipGetLastError(), "pcg iteration");
    };
    auto stopped = [&] {
        hipCheck(hipMemcpyAsync(c.mailbox, w.fl.p, sizeof(int) * F_COUNT, hipMemcpyDeviceToHost, c.stream), "poll");
        hipCheck(hipStreamSynchronize(c.stream), "poll sync");
        return reinterpret_cast<const int*>(c.mailbox)[F_STOP] != 0;
    };
    if (pc.graph && !c.profiling && c.stream != nullptr)
    {
        // A group of iterations captured once and replayed between polls. The group starts on an even
        // iteration, so the direction buffers alternate as in the enqueued loop; the device stops every
        // kernel at convergence or maxIter, so the iterations of a group past either do nothing.
        const int group = (std::max(1, pc.checkEvery) + 1) & ~1;
        std::vector<uint64_t> key = {amgLayoutKey(a), reinterpret_cast<uintptr_t>(c.stream),
                                     reinterpret_cast<uintptr_t>(psi), reinterpret_cast<uintptr_t>(diag),
                                     reinterpret_cast<uintptr_t>(c.partials.p), reinterpret_cast<uintptr_t>(w.r.p),
                                     reinterpret_cast<uintptr_t>(w.p0.p), reinterpret_cast<uintptr_t>(w.p1.p),
                                     reinterpret_cast<uintptr_t>(w.q.p), reinterpret_cast<uintptr_t>(w.rOld.p),
                                     reinterpret_cast<uintptr_t>(w.st.p), reinterpret_cast<uintptr_t>(w.fl.p),
                                     reinterpret_cast<uintptr_t>(sA), reinterpret_cast<uintptr_t>(sB),
                                     (uint64_t)group, (uint64_t)pc.flexible, (uint64_t)pc.nPre, (uint64_t)pc.nPost,
                                     (uint64_t)pc.scaleCorrection, (uint64_t)seq};
        if (!w.graph || w.graphKey != key)
        {
            if (w.graph) (void)hipGraphExecDestroy(w.graph);
            w.graph = nullptr;
            hipGraph_t gr = nullptr;
            hipCheck(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal), "capture begin");
            try
            {
                for (int j = 0; j < group; ++j) iteration(j);
            }
            catch (...)
            {
                (void)hipStreamEndCapture(c.stream, &gr);
                if (gr) (void)hipGraphDestroy(gr);
                throw;
            }
            hipCheck(hipStreamEndCapture(c.stream, &gr), "capture end");
            const hipError_t e = hipGraphInstantiate(&w.graph, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            hipCheck(e, "graph instantiate");
            w.graphKey = key;
        }
        for (int it = 0; it < pc.maxIter; it += group)
        {
            hipCheck(hipGraphLaunch(w.graph, c.stream), "graph launch");
            if (stopped()) break;
        }
    }
    else
    {
        for (int it = 0; it < pc.maxIter; ++it)
        {
            iteration(it);
            if (((it + 1) % std::max(1, pc.checkEvery) == 0 || it + 1 == pc.maxIter) && stopped()) break;
        }
    }
    hipCheck(hipMemcpyAsync(c.mailbox, w.st.p, sizeof(double) * S_COUNT, hipMemcpyDeviceToHost, c.stream), "report");
    hipCheck(hipMemcpyAsync(c.mailbox + S_COUNT, w.fl.p, sizeof(int) * F_COUNT, hipMemcpyDeviceToHost, c.stream),
             "report flags");
    hipCheck(hipStreamSynchronize(c.stream), "report sync");
    const double* st = c.mailbox;
    const int* fl = reinterpret_cast<const int*>(c.mailbox + S_COUNT);
    PcgPerf perf;
    perf.initialResidual = st[S_INIT];
    perf.finalResidual = st[S_FINAL];
    perf.nIterations = fl[F_ITER];
    perf.converged = fl[F_CONV] != 0;
    return perf;
}

} // namespace

PcgPerf amgPcgSolve(Ctx& c, const Mesh& m, Amg& a, const double* diag, const double* upper, const double* b,
                    double* psi, const double* normFactor, const PcgControls& pc)
{
    if (!a.galerkinDone) throw Error(kErrState, "mfv: amg_galerkin must run before the solve");
    if (a.prec != pc.prec) throw Error(kErrState, "mfv: the hierarchy was prepared in another precision");
    if (a.nFine != m.nCells) invalid("pcg: the hierarchy was built for another mesh");
    if (pc.maxIter < 0 || pc.minIter < 0) invalid("pcg: negative iteration limits");
    (void)upper;
    if (!a.work) a.work.reset(new PcgWork());
    PcgWork& w = *a.work;
    const size_t n = (size_t)m.nCells;
    if (w.r.n != n)
    {
        w.r.resize(n);
        w.p0.resize(n);
        w.p1.resize(n);
        w.q.resize(n);
        w.rOld.resize(n);
        w.st.resize(S_COUNT);
        w.fl.resize(F_COUNT);
        hipCheck(hipMemsetAsync(w.rOld.p, 0, n * sizeof(double), c.stream), "pcg work");
    }
    Grid& g0 = a.grids[0];
    if (pc.prec == Prec::F64)
        return solveT<double, double>(c, m, a, diag, b, psi, normFactor, pc, g0.b64.p, g0.x64.p, g0.diagV64.p);
    if (pc.prec == Prec::F32)
        return solveT<float, float>(c, m, a, diag, b, psi, normFactor, pc, g0.b32.p, g0.x32.p, g0.diagV32.p);
    return solveT<float, __half>(c, m, a, diag, b, psi, normFactor, pc, g0.b16.p, g0.x16.p, g0.diagV32.p);
}

} // namespace mfv
