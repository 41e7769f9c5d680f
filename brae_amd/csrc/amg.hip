// AMG hierarchy: sliced-row layout build, Galerkin coarse matrices, and the V-cycle kernels in FP64,
// FP32 and FP16 (scaled) precision.
//
// Reference semantics (brae's device path): Galerkin gathers in ascending fine index
// (device_amg.cu:156-223, 238-284); V(nPre, nPost) weighted-Jacobi cycle with omega 0.8 and the
// safeDiag floor, first pre-sweep from zero, injection restriction/prolongation, exact coarsest solve
// (device_amg_vcycle.cu:982-1143 FP64, :1658-1766 FP32; amg_kernels.cuh:17-64, 257-282); the
// coarsest grid solved with a precomputed dense inverse (the reference's default coarse kind,
// device_amg_coarse_default.cuh:13) computed here by Gauss-Jordan with partial pivoting and the
// reference's pivot floor 1e-14*max|diag| (device_amg_detail.cuh:93-94).
#include "amg.hpp"
#include "reduce.hpp"
#include "sell_device.hpp"

#include <algorithm>
#include <cstring>
#include <cmath>
#include <map>
#include <numeric>
#include <type_traits>

namespace mfv {

namespace {

constexpr int kTpb = 256;
constexpr int kMaxDenseCoarsest = 1024;   // the dense inverse's column buffer in LDS
constexpr int kAlphaInBlockParts = 1024;  // corrAmulK partials a corrSmoothK workgroup reduces itself
inline int blocks(long long n) { return (int)((n + kTpb - 1) / kTpb); }

// ---------------- V-cycle kernels ----------------
//
// Every kernel takes its grid's vectors as stored (Ts: the arithmetic type Tv, or FP16 for the FP16
// cycle) with `xm`, the grid's multiplier of solution-type vectors (null outside the FP16 cycle); see
// CycVec.

// r = b - A x  (amgSpmvF then residualT)
template <typename Tm, typename Tv, typename Ts>
__global__ __launch_bounds__(kTpb) void residualK(SellView S, const Tm* __restrict__ val, const Tv* __restrict__ diag,
                                                  const double* __restrict__ scale, const Ts* __restrict__ x,
                                                  const Ts* __restrict__ b, Ts* __restrict__ r,
                                                  const float* __restrict__ xm, const int* __restrict__ stop)
{
    if (stop && *stop) return;
    const int c = xcdBlock() * kTpb + threadIdx.x;
    if (c >= S.n) return;
    const CycVec<Tv, Ts> X(x, xm), B(b, nullptr), R(r, nullptr);
    const Tv ax = rowAx<Tm, Tv>(S, c, val, diag, scale, X);
    R.set(c, B(c) - ax);
}

// The coarse-to-fine injection, fused into whichever kernel finishes a coarse grid's cycle: coarse cell
// C's final value goes to every fine cell of its agglomerate (the fine grid's cellStart/cellList), so
// the fine grid reads its correction e = xc[map] as a plain vector (prolongT's values, written once).
template <typename Tv, typename Ts>
struct Scatter
{
    const int* __restrict__ start;
    const int* __restrict__ list;
    Ts* __restrict__ e;        // null: no fine grid waits for this one (grid 0)
    const float* xm;           // the fine grid's multiplier
    const int4* quads;         // the fine grid's agglomerates of up to four (may be null: the lists alone)
    __device__ void operator()(int C, Tv v) const
    {
        if (!e) return;
        const CycVec<Tv, Ts> E(e, xm);
        if (quads)
        {
            const int4 q = quads[C];
            if (q.x >= 0)
            {
                E.set(q.x, v);
                if (q.y >= 0) E.set(q.y, v);
                if (q.z >= 0) E.set(q.z, v);
                if (q.w >= 0) E.set(q.w, v);
                return;
            }
        }
        for (int k = start[C]; k < start[C + 1]; ++k) E.set(list[k], v);
    }
};

// xOut = x + omega*(b - A x)/D  (amgSpmvF then smoothT)
template <typename Tm, typename Tv, typename Ts>
__global__ __launch_bounds__(kTpb) void smoothK(SellView S, const Tm* __restrict__ val, const Tv* __restrict__ diag,
                                                const double* __restrict__ scale, const Ts* __restrict__ x,
                                                const Ts* __restrict__ b, Ts* __restrict__ xOut,
                                                const float* __restrict__ xm, Scatter<Tv, Ts> sc,
                                                const int* __restrict__ stop)
{
    if (stop && *stop) return;
    const int c = xcdBlock() * kTpb + threadIdx.x;
    if (c >= S.n) return;
    const CycVec<Tv, Ts> X(x, xm), B(b, nullptr), XO(xOut, xm);
    Tv dc;
    const Tv ax = rowAx<Tm, Tv>(S, c, val, diag, scale, X, &dc);
    const Tv xo = X(c) + Tv(kOmega) * (B(c) - ax) / safeDiag(dc);
    XO.set(c, xo);
    sc(c, xo);
}

// x + e at any cell: the prolonged iterate a fused first post-sweep works on
template <typename Tv, typename Ts>
struct SumX
{
    CycVec<Tv, Ts> x, e;
    __device__ Tv operator()(int j) const { return x(j) + e(j); }
};

// prolongation fused into the first post-sweep: xin = x + e (prolongT), then
// xOut = xin + omega*(b - A xin)/D
template <typename Tm, typename Tv, typename Ts>
__global__ __launch_bounds__(kTpb) void prolongSmoothK(SellView S, const Tm* __restrict__ val,
                                                       const Tv* __restrict__ diag, const double* __restrict__ scale,
                                                       const Ts* __restrict__ e, const Ts* __restrict__ x,
                                                       const Ts* __restrict__ b, Ts* __restrict__ xOut,
                                                       const float* __restrict__ xm, Scatter<Tv, Ts> sc,
                                                       const int* __restrict__ stop)
{
    if (stop && *stop) return;
    const int c = xcdBlock() * kTpb + threadIdx.x;
    if (c >= S.n) return;
    const SumX<Tv, Ts> X{CycVec<Tv, Ts>(x, xm), CycVec<Tv, Ts>(e, xm)};
    const CycVec<Tv, Ts> B(b, nullptr), XO(xOut, xm);
    Tv dc;
    const Tv ax = rowAx<Tm, Tv>(S, c, val, diag, scale, X, &dc);
    const Tv xo = X(c) + Tv(kOmega) * (B(c) - ax) / safeDiag(dc);
    XO.set(c, xo);
    sc(c, xo);
}

// Scaled coarse correction, first half: e = xc[map] (prolonged), Ae = A e on the grid (e gathered at the
// neighbours), and the two sums r.e, e.Ae accumulated in FP64; the last workgroup forms
// alpha = (r.e)/(e.Ae) with the reference's guard (scaleFactorK: 1 when e.Ae <= 1e-30).
struct AlphaPost
{
    double* alpha;
    __device__ void operator()(double num, double den) const { *alpha = den > 1e-30 ? num / den : 1.0; }
};

template <typename Tm, typename Tv, typename Ts>
__global__ __launch_bounds__(kReduceBlock) void corrAmulK(SellView S, const Tm* __restrict__ val,
                                                          const Tv* __restrict__ diag, const double* __restrict__ scale,
                                                          const Ts* __restrict__ e, const Ts* __restrict__ r,
                                                          Ts* __restrict__ Ae, const float* __restrict__ xm, double* pa,
                                                          double* pb, double* seqA, double* seqB,
                                                          const int* __restrict__ stop)
{
    if (stop && *stop) return;
    __shared__ double lds[kReduceBlock / 64];
    const CycVec<Tv, Ts> E(e, xm), R(r, nullptr), AE(Ae, nullptr);
    double num = 0.0, den = 0.0;
    MFV_ROWS(S.n, c)
    {
        const Tv ae = rowAx<Tm, Tv>(S, c, val, diag, scale, E);
        AE.set(c, ae);
        const double ev = (double)E(c);
        const double tn = ev * (double)R(c), td = ev * (double)ae;
        if (seqA)
        {
            seqA[c] = tn;
            seqB[c] = td;
        }
        else
        {
            num += tn;
            den += td;
        }
    }
    if (seqA) return;
    const double sn = blockSum(num, lds);
    const double sd = blockSum(den, lds);
    if (threadIdx.x == 0)
    {
        pa[blockIdx.x] = sn;
        pb[blockIdx.x] = sd;
    }
}

// second half, in place: x = x + alpha*e + omega*(r - alpha*Ae)/D (smooth) or x + alpha*e (no sweep).
// With nParts > 0 every workgroup finishes alpha itself from corrAmulK's partials pa, pb (finish2K's
// arithmetic, so the same bits) and no finishing launch runs in between; else alpha is read.
template <typename Tv, typename Ts>
__global__ __launch_bounds__(kTpb) void corrSmoothK(int n, const Ts* __restrict__ e, const Ts* __restrict__ r,
                                                    const Ts* __restrict__ Ae, const Tv* __restrict__ diag,
                                                    const double* __restrict__ alpha, int smooth, Ts* __restrict__ x,
                                                    const float* __restrict__ xm, Scatter<Tv, Ts> sc,
                                                    const int* __restrict__ stop, int nParts,
                                                    const double* __restrict__ pa, const double* __restrict__ pb)
{
    static_assert(kTpb == kReduceBlock, "alpha is finished by kReduceBlock-thread workgroups");
    if (stop && *stop) return;
    __shared__ double lds[2 * (kFinishBlock / 64)];
    __shared__ double shAlpha;
    double alv;
    if (nParts > 0)
    {
        double num, den;
        finish2InBlock(nParts, pa, pb, lds, num, den);
        if (threadIdx.x == 0) shAlpha = den > 1e-30 ? num / den : 1.0;
        __syncthreads();
        alv = shAlpha;
    }
    else
        alv = *alpha;
    const int c = xcdBlock() * kTpb + threadIdx.x;
    if (c >= n) return;
    const CycVec<Tv, Ts> E(e, xm), R(r, nullptr), AE(Ae, nullptr), X(x, xm);
    const Tv al = (Tv)alv;
    const Tv ev = E(c);
    const Tv xo = smooth ? X(c) + al * ev + Tv(kOmega) * (R(c) - al * AE(c)) / safeDiag(diag[c]) : X(c) + al * ev;
    X.set(c, xo);
    sc(c, xo);
}

// The same for the FP16 cycle, four rows a thread: 8-byte loads of four FP16 values (and 16 bytes of the
// FP32 diagonal) instead of 2-byte ones, the arithmetic row by row as above (grid 0, which scatters nothing)
__global__ __launch_bounds__(kTpb) void corrSmoothH4K(int n, const __half* __restrict__ e, const __half* __restrict__ r,
                                                      const __half* __restrict__ Ae, const float* __restrict__ diag,
                                                      const double* __restrict__ alpha, int smooth,
                                                      __half* __restrict__ x, const float* __restrict__ xm,
                                                      Scatter<float, __half> sc, const int* __restrict__ stop)
{
    if (stop && *stop) return;
    const int c0 = 4 * (xcdBlock() * kTpb + threadIdx.x);
    if (c0 >= n) return;
    const CycVec<float, __half> E(e, xm), R(r, nullptr), AE(Ae, nullptr), X(x, xm);
    const float al = (float)(*alpha);
    if (c0 + 4 > n)
    {
        for (int c = c0; c < n; ++c)
        {
            const float ev = E(c);
            const float xo = smooth ? X(c) + al * ev + float(kOmega) * (R(c) - al * AE(c)) / safeDiag(diag[c])
                                    : X(c) + al * ev;
            X.set(c, xo);
            sc(c, xo);
        }
        return;
    }
    union H4
    {
        uint2 u;
        __half h[4];
    };
    H4 he, hr, ha, hx;
    he.u = *reinterpret_cast<const uint2*>(e + c0);
    hr.u = *reinterpret_cast<const uint2*>(r + c0);
    ha.u = *reinterpret_cast<const uint2*>(Ae + c0);
    hx.u = *reinterpret_cast<const uint2*>(x + c0);
    const float4 d4 = ldStream(reinterpret_cast<const float4*>(diag + c0));
    const float dd[4] = {d4.x, d4.y, d4.z, d4.w};
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
        const float ev = __half2float(he.h[k]) * E.im, xv = __half2float(hx.h[k]) * X.im;
        const float xo = smooth ? xv + al * ev + float(kOmega) * (__half2float(hr.h[k]) - al * __half2float(ha.h[k])) /
                                                      safeDiag(dd[k])
                                : xv + al * ev;
        hx.h[k] = __float2half(fminf(fmaxf(xo * X.m, -65504.0f), 65504.0f));
        sc(c0 + k, xo);
    }
    *reinterpret_cast<uint2*>(x + c0) = hx.u;
}

// prolongation alone (no post-sweep): x += e
template <typename Tv, typename Ts>
__global__ __launch_bounds__(kTpb) void prolongK(int n, const Ts* __restrict__ e, Ts* __restrict__ x,
                                                 const float* __restrict__ xm, Scatter<Tv, Ts> sc,
                                                 const int* __restrict__ stop)
{
    if (stop && *stop) return;
    const int c = xcdBlock() * kTpb + threadIdx.x;
    if (c >= n) return;
    const CycVec<Tv, Ts> E(e, xm), X(x, xm);
    const Tv xo = X(c) + E(c);
    X.set(c, xo);
    sc(c, xo);
}

// coarse rhs = sum of the fine residual over each agglomerate, ascending fine cell (restrictGatherT);
// with fromZero the coarse grid's first sweep from zero x = 0 + omega*(b - 0)/D (smoothFromZeroT)
// (FromZero a template parameter: its diagonal load then issues with the children's, not behind a
// branch; without it diagC may be null -- the coarsest grid.)
template <typename Tv, typename Ts, bool FromZero>
__global__ __launch_bounds__(kTpb) void restrictK(int nc, const int* __restrict__ cellStart,
                                                  const int* __restrict__ cellList, const int4* __restrict__ quads,
                                                  const Ts* __restrict__ r, Ts* __restrict__ bc,
                                                  const Tv* __restrict__ diagC, Ts* __restrict__ xc,
                                                  const float* __restrict__ xmC, const int* __restrict__ stop)
{
    if (stop && *stop) return;
    const int C = xcdBlock() * kTpb + threadIdx.x;
    if (C >= nc) return;
    const CycVec<Tv, Ts> R(r, nullptr), BC(bc, nullptr), XC(xc, xmC);
    Tv s = Tv(0);
    const int4 q = quads[C];
    const Tv dC = FromZero ? diagC[C] : Tv(1);
    if (q.x >= 0)
    {
        // the four children's loads issue together (an absent child reads the first one), and only the
        // present ones are summed, in ascending order
        const Tv v0 = R(q.x), v1 = R(q.y >= 0 ? q.y : q.x), v2 = R(q.z >= 0 ? q.z : q.x),
                 v3 = R(q.w >= 0 ? q.w : q.x);
        s += v0;
        s = q.y >= 0 ? s + v1 : s;
        s = q.z >= 0 ? s + v2 : s;
        s = q.w >= 0 ? s + v3 : s;
    }
    else
        for (int k = cellStart[C]; k < cellStart[C + 1]; ++k) s += R(cellList[k]);
    BC.set(C, s);
    if (FromZero) XC.set(C, Tv(0) + Tv(kOmega) * (s - Tv(0)) / safeDiag(dC));
}

template <typename Ts>
__global__ __launch_bounds__(kTpb) void zeroK(int n, Ts* __restrict__ x, const int* __restrict__ stop)
{
    if (stop && *stop) return;
    const int c = xcdBlock() * kTpb + threadIdx.x;
    if (c < n) x[c] = Ts(0.0f);
}

// grid 0 entry from the FP64 residual: b = (Tv) r (the FP16 cycle: r over its power-of-two scale sig),
// x = omega*b/D when pre-smoothing from zero
template <typename Tv, typename Ts>
__global__ __launch_bounds__(kTpb) void startK(int n, const double* __restrict__ r, const Tv* __restrict__ diag,
                                               Ts* __restrict__ b, Ts* __restrict__ x, const float* __restrict__ xm,
                                               const double* __restrict__ sig, int fromZero,
                                               const int* __restrict__ stop)
{
    if (stop && *stop) return;
    const int c = xcdBlock() * kTpb + threadIdx.x;
    if (c >= n) return;
    const CycVec<Tv, Ts> B(b, nullptr), X(x, xm);
    const Tv bv = sig ? (Tv)(r[c] * (1.0 / *sig)) : (Tv)r[c];
    B.set(c, bv);
    X.set(c, fromZero ? Tv(0) + Tv(kOmega) * (bv - Tv(0)) / safeDiag(diag[c]) : Tv(0));
}

// sum_j row[j] * v[j] in ascending j (the dense coarsest apply's row), eight loads in flight ahead of
// the adds
__device__ __forceinline__ double denseRowDot(const double* row, const double* v, int n)
{
    double s = 0.0;
    int j = 0;
    for (; j + 8 <= n; j += 8)
    {
        double a[8], b[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
        {
            a[u] = row[j + u];
            b[u] = v[j + u];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) s += a[u] * b[u];
    }
    for (; j < n; ++j) s += row[j] * v[j];
    return s;
}

// coarsest grid: x = Ainv * b in FP64, ascending column per row (one workgroup)
template <typename Tv, typename Ts>
__global__ __launch_bounds__(kTpb) void coarseApplyK(int n, const double* __restrict__ inv, const Ts* __restrict__ b,
                                                     Ts* __restrict__ x, const float* __restrict__ xm,
                                                     Scatter<Tv, Ts> sc, const int* __restrict__ stop)
{
    if (stop && *stop) return;
    __shared__ double sb[kMaxDenseCoarsest];
    const CycVec<Tv, Ts> B(b, nullptr), X(x, xm);
    for (int j = threadIdx.x; j < n; j += blockDim.x) sb[j] = (double)B(j);
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x)
    {
        const double s = denseRowDot(inv + (size_t)i * n, sb, n);
        X.set(i, (Tv)s);
        sc(i, (Tv)s);
    }
}

// max and min of |d| over a workgroup's rows, for the FP16 cycle's solution-vector multiplier
__global__ __launch_bounds__(kReduceBlock) void diagRangeK(int n, const double* __restrict__ d, double* pmax,
                                                           double* pmin)
{
    __shared__ double lds[kReduceBlock / 64];
    double mx = 0.0, mn = 1e308;
    MFV_ROWS(n, i)
    {
        const double v = fabs(d[i]);
        mx = fmax(mx, v);
        if (v > 0.0) mn = fmin(mn, v);
    }
    const double bmx = blockMax(mx, lds);
    const double bmn = -blockMax(-mn, lds);
    if (threadIdx.x == 0)
    {
        pmax[blockIdx.x] = bmx;
        pmin[blockIdx.x] = bmn;
    }
}

// the grid's multiplier: the power of two at or below the geometric mean of max|d| and min|d| (the
// solution vectors, about b/D, then sit near 1 whatever the units)
__global__ __launch_bounds__(kFinishBlock) void xmulFinishK(int nParts, const double* __restrict__ pmax,
                                                            const double* __restrict__ pmin, float* __restrict__ out)
{
    __shared__ double lds[kFinishBlock / 64];
    double mx = finishMax(nParts, pmax, 0.0);
    double mn = finishMax<true>(nParts, pmin, 1e308);
    mx = blockMax<kFinishBlock>(mx, lds);
    mn = -blockMax<kFinishBlock>(-mn, lds);
    if (threadIdx.x) return;
    if (!(mx > 0.0)) mx = 1.0;
    if (!(mn > 0.0) || mn > mx) mn = mx;
    *out = (float)exp2(floor(0.5 * (log2(mx) + log2(mn))));
}

// ---------------- the small grids in one workgroup ----------------
//
// Below kTailCells cells a grid's kernels are launch-bound (a few microseconds each, whatever their
// size). The tail kernel runs the whole cycle from grid tailStart down to the coarsest and back in one
// workgroup of kTailThreads threads, with a barrier between steps: the same steps, in the same order
// and arithmetic as the per-grid kernels, the scaled correction's two sums reduced over the workgroup
// in a fixed order. It hands its first grid's values up through the same injection as the per-grid
// path. (The sequential-reduction mode keeps the per-grid kernels, whose sums it defines.)
constexpr int kTailThreads = 1024;
constexpr int kTailCells = 4096;
constexpr int kMaxTail = 32;

struct TailGrid
{
    SellView S;
    const void* val;
    const void* diag;
    const double* scale;
    void *b, *x, *y, *r, *e;
    const float* xm;        // solution-vector multiplier (FP16 cycle), else null
    const int* cellStart;   // this grid's cells per coarse cell of the next grid
    const int* cellList;
    int n;
};

template <typename Tm, typename Tv, typename Ts>
__device__ __forceinline__ void tailSweep(const TailGrid& G, Ts* xin, Ts* xout, const Scatter<Tv, Ts>& sc)
{
    const Tm* val = static_cast<const Tm*>(G.val);
    const Tv* diag = static_cast<const Tv*>(G.diag);
    const CycVec<Tv, Ts> X(xin, G.xm), XO(xout, G.xm), B(static_cast<Ts*>(G.b), nullptr);
    for (int c = threadIdx.x; c < G.n; c += kTailThreads)
    {
        const Tv ax = rowAx<Tm, Tv>(G.S, c, val, diag, G.scale, X);
        const Tv xo = X(c) + Tv(kOmega) * (B(c) - ax) / safeDiag(diag[c]);
        XO.set(c, xo);
        sc(c, xo);
    }
}

template <typename Tm, typename Tv, typename Ts>
__global__ __launch_bounds__(kTailThreads) void tailK(const TailGrid* __restrict__ D, int T,
                                                     const double* __restrict__ inv, int nPre, int nPost,
                                                     int scaled, Scatter<Tv, Ts> up, const int* __restrict__ stop)
{
    if (stop && *stop) return;
    __shared__ double lds[kTailThreads / 64];
    __shared__ double sb[1024];
    __shared__ double shAlpha;
    const int tid = threadIdx.x;
    const Scatter<Tv, Ts> none{nullptr, nullptr, nullptr, nullptr, nullptr};
    // grid g's current iterate is its x buffer, or y after an odd number of ping-pong sweeps (bit g)
    unsigned flip = 0;
    auto X = [&](int g) { return static_cast<Ts*>((flip >> g) & 1u ? D[g].y : D[g].x); };
    auto Y = [&](int g) { return static_cast<Ts*>((flip >> g) & 1u ? D[g].x : D[g].y); };
    // the finishing step of grid g hands its values to grid g-1's e, or out of the tail for g = 0
    auto upOf = [&](int g) {
        if (g == 0) return up;
        return Scatter<Tv, Ts>{D[g - 1].cellStart, D[g - 1].cellList, static_cast<Ts*>(D[g - 1].e), D[g - 1].xm,
                               nullptr};
    };
    // ---- down ----
    for (int g = 0; g + 1 < T; ++g)
    {
        const TailGrid G = D[g];
        const Tm* val = static_cast<const Tm*>(G.val);
        const Tv* diag = static_cast<const Tv*>(G.diag);
        const CycVec<Tv, Ts> B(static_cast<Ts*>(G.b), nullptr), R(static_cast<Ts*>(G.r), nullptr);
        for (int s = 1; s < nPre; ++s)
        {
            tailSweep<Tm, Tv, Ts>(G, X(g), Y(g), none);
            __syncthreads();
            flip ^= 1u << g;
        }
        const CycVec<Tv, Ts> Xg(X(g), G.xm);
        if (nPre == 0)
        {
            for (int c = tid; c < G.n; c += kTailThreads) Xg.set(c, Tv(0));
            __syncthreads();
        }
        for (int c = tid; c < G.n; c += kTailThreads) R.set(c, B(c) - rowAx<Tm, Tv>(G.S, c, val, diag, G.scale, Xg));
        __syncthreads();
        const TailGrid C = D[g + 1];
        const Tv* dC = static_cast<const Tv*>(C.diag);
        const CycVec<Tv, Ts> BC(static_cast<Ts*>(C.b), nullptr), XC(X(g + 1), C.xm);
        const int fromZero = (g + 2 < T && nPre > 0) ? 1 : 0;
        for (int k = tid; k < C.n; k += kTailThreads)
        {
            Tv s = Tv(0);
            for (int q = G.cellStart[k]; q < G.cellStart[k + 1]; ++q) s += R(G.cellList[q]);
            BC.set(k, s);
            if (fromZero) XC.set(k, Tv(0) + Tv(kOmega) * (s - Tv(0)) / safeDiag(dC[k]));
        }
        __syncthreads();
    }
    // ---- coarsest: x = Ainv b ----
    {
        const TailGrid G = D[T - 1];
        const CycVec<Tv, Ts> B(static_cast<Ts*>(G.b), nullptr), Xg(X(T - 1), G.xm);
        for (int j = tid; j < G.n; j += kTailThreads) sb[j] = (double)B(j);
        __syncthreads();
        const Scatter<Tv, Ts> sc = upOf(T - 1);
        for (int i = tid; i < G.n; i += kTailThreads)
        {
            const double s = denseRowDot(inv + (size_t)i * G.n, sb, G.n);
            Xg.set(i, (Tv)s);
            sc(i, (Tv)s);
        }
        __syncthreads();
    }
    // ---- up ----
    const int nExtra = nPost > 1 ? nPost - 1 : 0;
    for (int g = T - 2; g >= 0; --g)
    {
        const TailGrid G = D[g];
        const Tm* val = static_cast<const Tm*>(G.val);
        const Tv* diag = static_cast<const Tv*>(G.diag);
        const CycVec<Tv, Ts> R(static_cast<Ts*>(G.r), nullptr), E(static_cast<Ts*>(G.e), G.xm);
        const Scatter<Tv, Ts> sc = upOf(g);
        if (scaled)
        {
            const CycVec<Tv, Ts> Xg(X(g), G.xm), AE(Y(g), nullptr);
            double num = 0.0, den = 0.0;
            for (int c = tid; c < G.n; c += kTailThreads)
            {
                const Tv ae = rowAx<Tm, Tv>(G.S, c, val, diag, G.scale, E);
                AE.set(c, ae);
                const double ev = (double)E(c);
                num += ev * (double)R(c);
                den += ev * (double)ae;
            }
            num = blockSum<kTailThreads>(num, lds);
            den = blockSum<kTailThreads>(den, lds);
            if (tid == 0) shAlpha = den > 1e-30 ? num / den : 1.0;
            __syncthreads();
            const Tv al = (Tv)shAlpha;
            for (int c = tid; c < G.n; c += kTailThreads)
            {
                const Tv ev = E(c);
                const Tv xo = nPost > 0 ? Xg(c) + al * ev + Tv(kOmega) * (R(c) - al * AE(c)) / safeDiag(diag[c])
                                        : Xg(c) + al * ev;
                Xg.set(c, xo);
                if (!nExtra) sc(c, xo);
            }
            __syncthreads();
        }
        else if (nPost > 0)
        {
            const CycVec<Tv, Ts> B(static_cast<Ts*>(G.b), nullptr), Yg(Y(g), G.xm);
            const SumX<Tv, Ts> XE{CycVec<Tv, Ts>(X(g), G.xm), E};
            for (int c = tid; c < G.n; c += kTailThreads)
            {
                const Tv ax = rowAx<Tm, Tv>(G.S, c, val, diag, G.scale, XE);
                const Tv xo = XE(c) + Tv(kOmega) * (B(c) - ax) / safeDiag(diag[c]);
                Yg.set(c, xo);
                if (!nExtra) sc(c, xo);
            }
            __syncthreads();
            flip ^= 1u << g;
        }
        else
        {
            const CycVec<Tv, Ts> Xg(X(g), G.xm);
            for (int c = tid; c < G.n; c += kTailThreads)
            {
                const Tv xo = Xg(c) + E(c);
                Xg.set(c, xo);
                sc(c, xo);
            }
            __syncthreads();
        }
        if (scaled || nPost > 0)
            for (int s = 0; s < nExtra; ++s)
            {
                tailSweep<Tm, Tv, Ts>(G, X(g), Y(g), s + 1 == nExtra ? sc : none);
                __syncthreads();
                flip ^= 1u << g;
            }
    }
}

// ---------------- the small grids resident in LDS ----------------
//
// tailK's cycle with the coarsest grids' arrays staged into the workgroup's LDS when the launch starts.
// A tail step is a chain of dependent loads (slice start, columns and coefficients, the gathered
// vector); from HBM that chain costs microseconds per step, from LDS a fraction of one. The tail's
// grids from TG on (the longest run of coarsest grids that fits the LDS a workgroup may hold) are
// staged: slice starts, 16-bit columns and 16-bit restriction lists from a blob built once per
// hierarchy and precision (the same byte offsets in the blob and in LDS), the coefficients and
// diagonals copied from the grids' own arrays at every launch (a Galerkin may have changed them), the
// vectors held in LDS only. The grids before TG are read where they are, as tailK reads them. Values
// cross into and out of the tail where tailK's do: the first grid's right-hand side and start iterate
// from its buffers, the injection into the finer grid's e through `up`. Each row's sum is rowAx's (entry
// order, padded slots included) and each step tailK's, so the result is the same bits. The scaled
// correction keeps tailK.

struct LdsTailGrid
{
    int n, nnz;
    int oSlice, oCols, oStart, oList;    // static arrays (blob and LDS byte offsets)
    int oVal, oDiag, oB, oX, oY, oR, oE; // staged and LDS-only arrays (LDS byte offsets)
};

// one contiguous run of 4-byte words copied from global memory into LDS when the launch starts
struct StageSeg
{
    const void* src;
    int dst;     // LDS byte offset
    int first;   // its first word in the launch's copy plan
    int words;
};
constexpr int kMaxStageSegs = 2 + 2 * kMaxTail;   // the blob, the inverse, each staged grid's values and diagonal

struct LdsTail
{
    int T, TG;         // the tail's grids; the first TG are read where they are
    int oSb;           // the coarsest right-hand side in FP64
    int oInv;          // the coarsest inverse (n*n FP64) staged, or -1: read where it is
    int bytes;         // dynamic LDS the launch needs
    int nSegs, words;  // the copy plan: segments and their total words
    StageSeg seg[kMaxStageSegs];
    LdsTailGrid g[kMaxTail];   // [t]: grid tailStart + t, for t >= TG
};
constexpr int kStageBatch = 16;   // words a thread loads before it stores any

// one tail grid as the kernel sees it: vectors, diagonal and coefficients where they are (global or
// LDS), its rows either as sliced rows in global memory or as the staged 16-bit form
struct MixGrid
{
    int n, staged;
    SellView S;
    const void* val;
    const void* diag;
    const int* sp;
    const unsigned short* cols;
    void *b, *x, *y, *r, *e;
    const int* cs32;
    const int* cl32;
    const unsigned short* cs16;
    const unsigned short* cl16;
    double scale;
    float m, im;
    __device__ int start(int k) const { return cs16 ? (int)cs16[k] : cs32[k]; }
    __device__ int list(int k) const { return cl16 ? (int)cl16[k] : cl32[k]; }
};

// A x at row c of a staged grid: rowAx's sum on the 16-bit columns
template <typename Tm, typename Tv, typename XF>
__device__ __forceinline__ Tv stagedRowAx(const MixGrid& G, int c, XF X, Tv* dOut)
{
    const Tm* val = static_cast<const Tm*>(G.val);
    const Tv xc = X(c);
    const Tv dc = static_cast<const Tv*>(G.diag)[c];
    *dOut = dc;
    const int s = c >> 6, lane = c & 63;
    const int p0 = G.sp[s], w = (G.sp[s + 1] - p0) >> 6;
    Tv acc = sizeof(Tm) == 2 ? Tv(0) : dc * xc;
    // eight entries' loads issued before their gathers, the sum in entry order (a coarse grid's rows
    // are tens of entries wide: one LDS round trip per entry would serialise the step)
    int k = 0;
    for (; k + 8 <= w; k += 8)
    {
        Tv a[8], xv[8];
        int col[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
        {
            const int e = p0 + ((k + u) << 6) + lane;
            a[u] = cvt<Tv>(val[e]);
            col[u] = G.cols[e];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) xv[u] = X(col[u]);
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += a[u] * xv[u];
    }
    for (; k < w; ++k)
    {
        const int e = p0 + (k << 6) + lane;
        acc += cvt<Tv>(val[e]) * X((int)G.cols[e]);
    }
    if constexpr (sizeof(Tm) == 2)
        return dc * xc + (Tv)G.scale * acc;
    else
        return acc;
}

template <typename Tm, typename Tv, typename XF>
__device__ __forceinline__ Tv mixRowAx(const MixGrid& G, int c, XF X, Tv* dOut)
{
    if (G.staged) return stagedRowAx<Tm, Tv>(G, c, X, dOut);
    return rowAx<Tm, Tv>(G.S, c, static_cast<const Tm*>(G.val), static_cast<const Tv*>(G.diag), &G.scale, X, dOut);
}

template <typename Tm, typename Tv, typename Ts>
__global__ __launch_bounds__(kTailThreads) void tailMixK(const TailGrid* __restrict__ Dg,
                                                        const LdsTail* __restrict__ Lg,
                                                        const double* __restrict__ inv, int nPre, int nPost,
                                                        Scatter<Tv, Ts> up, const int* __restrict__ stop)
{
    if (stop && *stop) return;
    extern __shared__ __align__(16) uint8_t L[];
    __shared__ MixGrid M[kMaxTail];
    __shared__ StageSeg sSeg[kMaxStageSegs];
    __shared__ int sInv, sSb, sSegs, sWords;
    const int tid = threadIdx.x;
    // the descriptors and the copy plan into LDS (one round of loads), then the copy itself with
    // kStageBatch loads in flight per thread: the staging costs a few memory round trips, not one per
    // array and chunk
    const int T = Lg->T, TG = Lg->TG;
    if (tid == 0)
    {
        sInv = Lg->oInv;
        sSb = Lg->oSb;
        sSegs = Lg->nSegs;
        sWords = Lg->words;
    }
    if (tid < Lg->nSegs) sSeg[tid] = Lg->seg[tid];
    if (tid < T)
    {
        const TailGrid& D = Dg[tid];
        MixGrid q;
        q.n = D.n;
        q.staged = tid >= TG;
        q.S = D.S;
        q.scale = *D.scale;
        q.m = D.xm ? *D.xm : 1.0f;
        q.im = D.xm ? 1.0f / *D.xm : 1.0f;
        if (!q.staged)
        {
            q.val = D.val;
            q.diag = D.diag;
            q.sp = nullptr;
            q.cols = nullptr;
            q.b = D.b; q.x = D.x; q.y = D.y; q.r = D.r; q.e = D.e;
            q.cs32 = D.cellStart;
            q.cl32 = D.cellList;
            q.cs16 = nullptr;
            q.cl16 = nullptr;
        }
        else
        {
            const LdsTailGrid& G = Lg->g[tid];
            q.val = L + G.oVal;
            q.diag = L + G.oDiag;
            q.sp = reinterpret_cast<const int*>(L + G.oSlice);
            q.cols = reinterpret_cast<const unsigned short*>(L + G.oCols);
            q.b = L + G.oB; q.x = L + G.oX; q.y = L + G.oY; q.r = L + G.oR; q.e = L + G.oE;
            q.cs32 = nullptr;
            q.cl32 = nullptr;
            q.cs16 = reinterpret_cast<const unsigned short*>(L + G.oStart);
            q.cl16 = reinterpret_cast<const unsigned short*>(L + G.oList);
        }
        M[tid] = q;
    }
    __syncthreads();
    {
        const int nSegs = sSegs, words = sWords;
        for (int base = tid; base < words; base += kStageBatch * kTailThreads)
        {
            unsigned v[kStageBatch];
            int at[kStageBatch];
            int sg = 0;
#pragma unroll
            for (int u = 0; u < kStageBatch; ++u)
            {
                const int w = base + u * kTailThreads;
                at[u] = -1;
                if (w < words)
                {
                    while (sg + 1 < nSegs && w >= sSeg[sg + 1].first) ++sg;
                    const int k = w - sSeg[sg].first;
                    v[u] = static_cast<const unsigned*>(sSeg[sg].src)[k];
                    at[u] = sSeg[sg].dst + 4 * k;
                }
            }
#pragma unroll
            for (int u = 0; u < kStageBatch; ++u)
                if (at[u] >= 0) *reinterpret_cast<unsigned*>(L + at[u]) = v[u];
        }
    }
    // the first grid's b and x when it is staged itself (the finer grid's restriction left them)
    if (TG == 0)
        for (int i = tid; i < M[0].n; i += kTailThreads)
        {
            const Ts bv = static_cast<const Ts*>(Dg[0].b)[i], xv = static_cast<const Ts*>(Dg[0].x)[i];
            static_cast<Ts*>(M[0].b)[i] = bv;
            static_cast<Ts*>(M[0].x)[i] = xv;
        }
    __syncthreads();
    const int nL = M[T - 1].n;
    const double* A = sInv >= 0 ? reinterpret_cast<const double*>(L + sInv) : inv;
    double* sb = reinterpret_cast<double*>(L + sSb);
    auto vec = [&](int g, void* p) { return CycVec<Tv, Ts>(static_cast<Ts*>(p), M[g].m, M[g].im); };
    auto rhs = [&](void* p) { return CycVec<Tv, Ts>(static_cast<Ts*>(p), nullptr); };
    unsigned flip = 0;
    auto X = [&](int g) { return (flip >> g) & 1u ? M[g].y : M[g].x; };
    auto Y = [&](int g) { return (flip >> g) & 1u ? M[g].x : M[g].y; };
    // grid g's finished value v at row c: into the finer grid's e, or out through `up` for grid 0
    auto inject = [&](int g, int c, Tv v) {
        if (g == 0)
        {
            up(c, v);
            return;
        }
        const MixGrid& F = M[g - 1];
        const CycVec<Tv, Ts> E = vec(g - 1, F.e);
        for (int k = F.start(c); k < F.start(c + 1); ++k) E.set(F.list(k), v);
    };
    // one Jacobi sweep of grid g: X -> Y, injecting the result when `last`
    auto sweep = [&](int g, bool last) {
        const MixGrid& G = M[g];
        const CycVec<Tv, Ts> Xg = vec(g, X(g)), Yg = vec(g, Y(g)), B = rhs(G.b);
        for (int c = tid; c < G.n; c += kTailThreads)
        {
            Tv dc;
            const Tv ax = mixRowAx<Tm, Tv>(G, c, Xg, &dc);
            const Tv xo = Xg(c) + Tv(kOmega) * (B(c) - ax) / safeDiag(dc);
            Yg.set(c, xo);
            if (last) inject(g, c, xo);
        }
        __syncthreads();
        flip ^= 1u << g;
    };
    // ---- down ----
    for (int g = 0; g + 1 < T; ++g)
    {
        const MixGrid& G = M[g];
        for (int s = 1; s < nPre; ++s) sweep(g, false);
        const CycVec<Tv, Ts> Xg = vec(g, X(g)), B = rhs(G.b), R = rhs(G.r);
        if (nPre == 0)
        {
            for (int c = tid; c < G.n; c += kTailThreads) Xg.set(c, Tv(0));
            __syncthreads();
        }
        for (int c = tid; c < G.n; c += kTailThreads)
        {
            Tv dc;
            R.set(c, B(c) - mixRowAx<Tm, Tv>(G, c, Xg, &dc));
        }
        __syncthreads();
        const MixGrid& C = M[g + 1];
        const Tv* dC = static_cast<const Tv*>(C.diag);
        const CycVec<Tv, Ts> BC = rhs(C.b), XC = vec(g + 1, X(g + 1));
        const int fromZero = (g + 2 < T && nPre > 0) ? 1 : 0;
        for (int k = tid; k < C.n; k += kTailThreads)
        {
            Tv s = Tv(0);
            for (int q = G.start(k); q < G.start(k + 1); ++q) s += R(G.list(q));
            BC.set(k, s);
            if (fromZero) XC.set(k, Tv(0) + Tv(kOmega) * (s - Tv(0)) / safeDiag(dC[k]));
        }
        __syncthreads();
    }
    // ---- coarsest: x = Ainv b ----
    {
        const MixGrid& G = M[T - 1];
        const CycVec<Tv, Ts> B = rhs(G.b), Xg = vec(T - 1, X(T - 1));
        for (int j = tid; j < G.n; j += kTailThreads) sb[j] = (double)B(j);
        __syncthreads();
        for (int i = tid; i < G.n; i += kTailThreads)
        {
            const double s = denseRowDot(A + (size_t)i * G.n, sb, G.n);
            Xg.set(i, (Tv)s);
            inject(T - 1, i, (Tv)s);
        }
        __syncthreads();
    }
    // ---- up ----
    const int nExtra = nPost > 1 ? nPost - 1 : 0;
    for (int g = T - 2; g >= 0; --g)
    {
        const MixGrid& G = M[g];
        const CycVec<Tv, Ts> E = vec(g, G.e);
        if (nPost > 0)
        {
            const CycVec<Tv, Ts> B = rhs(G.b), Yg = vec(g, Y(g));
            const SumX<Tv, Ts> XE{vec(g, X(g)), E};
            for (int c = tid; c < G.n; c += kTailThreads)
            {
                Tv dc;
                const Tv ax = mixRowAx<Tm, Tv>(G, c, XE, &dc);
                const Tv xo = XE(c) + Tv(kOmega) * (B(c) - ax) / safeDiag(dc);
                Yg.set(c, xo);
                if (!nExtra) inject(g, c, xo);
            }
            __syncthreads();
            flip ^= 1u << g;
            for (int s = 0; s < nExtra; ++s) sweep(g, s + 1 == nExtra);
        }
        else
        {
            const CycVec<Tv, Ts> Xg = vec(g, X(g));
            for (int c = tid; c < G.n; c += kTailThreads)
            {
                const Tv xo = Xg(c) + E(c);
                Xg.set(c, xo);
                inject(g, c, xo);
            }
            __syncthreads();
        }
    }
}

// ---------------- small solves batched into one launch ----------------
//
// A solve of a few tens of thousands of rows is latency-bound on the per-grid kernels: each of the
// thirty-odd dependent launches of a PCG iteration costs microseconds whatever its size, and the device
// sits mostly idle. batchPcgK runs whole solves, one workgroup each, many in one launch: every workgroup
// takes one job -- the flexible AMG-PCG of mfv_amg_pcg_solve (OpenFOAM's normFactor and residual, the
// Polak-Ribiere+ beta, the V-cycle in the job's precision with its FP16 input scale and, when asked,
// the scaled correction) -- and
// iterates it to its tolerance on its own, the steps separated by workgroup barriers and every sum
// reduced over the workgroup in a fixed order. A job's result depends on nothing else in the launch, so
// it is the same bits in any batch. No workgroup waits for another.
struct BatchJob
{
    int n, T;                  // rows; grids of the hierarchy
    const TailGrid* grids;     // [T], grid 0 first (as sliced rows), in the cycle's precision
    SellView S0;               // grid 0's rows for the FP64 operator
    const double* val0;        // its FP64 coefficients
    const double* diag;        // the folded FP64 diagonal
    const double* b;           // right-hand side
    double* x;                 // solution: initial guess in, result out
    double *r0, *r1, *p, *q;   // work vectors, n each
    const double* inv;         // the coarsest grid's dense inverse
    double tol, relTol;
    int maxIter, minIter, nPre, nPost, half, flexible, scaled;
    double* outD;              // [2]: initial and final residual
    int* outI;                 // [2]: iterations, converged
};

// threads per batched solve: 512 leaves each 256 registers (the cycle's row sums and the FP64 operator
// are live at once), 1024 spilled
constexpr int kBatchThreads = 512;

template <typename Tm, typename Tv, typename Ts>
__global__ __launch_bounds__(kBatchThreads) void batchPcgK(const BatchJob* __restrict__ jobs)
{
    __shared__ BatchJob J;
    __shared__ MixGrid M[kMaxTail];
    __shared__ double lds[kBatchThreads / 64];
    __shared__ double sb[kMaxDenseCoarsest];
    __shared__ double shA, shB;
    const int tid = threadIdx.x;
    static_assert(sizeof(BatchJob) % 4 == 0, "job copied as words");
    for (int i = tid; i < (int)(sizeof(BatchJob) / 4); i += kBatchThreads)
        reinterpret_cast<int*>(&J)[i] = reinterpret_cast<const int*>(jobs + blockIdx.x)[i];
    __syncthreads();
    const int n = J.n, T = J.T;
    if (tid < T)
    {
        const TailGrid& D = J.grids[tid];
        MixGrid q;
        q.n = D.n;
        q.staged = 0;
        q.S = D.S;
        q.scale = *D.scale;
        q.m = D.xm ? *D.xm : 1.0f;
        q.im = D.xm ? 1.0f / *D.xm : 1.0f;
        q.val = D.val;
        q.diag = D.diag;
        q.sp = nullptr;
        q.cols = nullptr;
        q.b = D.b; q.x = D.x; q.y = D.y; q.r = D.r; q.e = D.e;
        q.cs32 = D.cellStart;
        q.cl32 = D.cellList;
        q.cs16 = nullptr;
        q.cl16 = nullptr;
        M[tid] = q;
    }
    __syncthreads();
    const double* A = J.val0;
    const double* Dg = J.diag;
    double* x = J.x;
    const double* b = J.b;
    auto vec = [&](int g, void* p) { return CycVec<Tv, Ts>(static_cast<Ts*>(p), M[g].m, M[g].im); };
    auto rhs = [&](void* p) { return CycVec<Tv, Ts>(static_cast<Ts*>(p), nullptr); };
    // ---- r = b - A x, OpenFOAM's normFactor about the mean of x, the initial residual ----
    double sx = 0.0;
    for (int c = tid; c < n; c += kBatchThreads) sx += x[c];
    const double xRef = blockSum<kBatchThreads>(sx, lds) ;
    if (tid == 0) shA = xRef / (double)n;
    __syncthreads();
    {
        const double xr = shA;
        double sr = 0.0, snf = 0.0;
        for (int c = tid; c < n; c += kBatchThreads)
        {
            const double ax = rowOffSum<double, double>(J.S0, c, A, PlainX<double>{x}, Dg[c] * x[c]);
            const double rc = b[c] - ax;
            J.r0[c] = rc;
            const double sa = rowOffSum<double, double>(J.S0, c, A, [](int) { return 1.0; }, Dg[c]);
            const double t = sa * xr;
            sr += fabs(rc);
            snf += fabs(ax - t) + fabs(b[c] - t);
        }
        const double a = blockSum<kBatchThreads>(sr, lds);
        const double nfs = blockSum<kBatchThreads>(snf, lds);
        if (tid == 0)
        {
            shA = a;
            shB = nfs + 1e-20;
        }
        __syncthreads();
    }
    const double nf = shB, init = shA / nf;
    auto converged = [&](double res) { return res < J.tol || (J.relTol > 0.0 && res < J.relTol * init); };
    double res = init, rhoOld = 1e300;
    int it = 0;
    bool conv = converged(init);
    bool stop = J.minIter <= 0 && conv;
    double* rCur = J.r0;    // this iteration's residual; rPrev the one before (the flexible beta's r_old)
    double* rPrev = J.r1;
    const int fromZero = J.nPre > 0 ? 1 : 0;
    while (!stop)
    {
        const double* r = rCur;
        const double* rOld = rPrev;
        // ---- the cycle's input: b0 = r (over its power-of-two scale in FP16), the first sweep from zero ----
        double sig = 0.0;
        if (J.half)
        {
            double mx = 0.0;
            for (int c = tid; c < n; c += kBatchThreads) mx = fmax(mx, fabs(r[c]));
            mx = blockMax<kBatchThreads>(mx, lds);
            if (tid == 0) shA = mx > 0.0 ? exp2(ceil(log2(mx))) : 1.0;
            __syncthreads();
            sig = shA;
        }
        {
            const MixGrid& G = M[0];
            const CycVec<Tv, Ts> B = rhs(G.b), X0 = vec(0, G.x);
            const Tv* d0 = static_cast<const Tv*>(G.diag);
            for (int c = tid; c < n; c += kBatchThreads)
            {
                const Tv bv = J.half ? (Tv)(r[c] * (1.0 / sig)) : (Tv)r[c];
                B.set(c, bv);
                X0.set(c, fromZero ? Tv(0) + Tv(kOmega) * (bv - Tv(0)) / safeDiag(d0[c]) : Tv(0));
            }
            __syncthreads();
        }
        // ---- one V-cycle from grid 0 (tailK's steps), its result in grid 0's current x ----
        unsigned flip = 0;
        auto X = [&](int g) { return (flip >> g) & 1u ? M[g].y : M[g].x; };
        auto Y = [&](int g) { return (flip >> g) & 1u ? M[g].x : M[g].y; };
        auto inject = [&](int g, int c, Tv v) {
            if (g == 0) return;
            const MixGrid& F = M[g - 1];
            const CycVec<Tv, Ts> E = vec(g - 1, F.e);
            for (int k = F.start(c); k < F.start(c + 1); ++k) E.set(F.list(k), v);
        };
        auto sweep = [&](int g, bool last) {
            const MixGrid& G = M[g];
            const CycVec<Tv, Ts> Xg = vec(g, X(g)), Yg = vec(g, Y(g)), B = rhs(G.b);
            for (int c = tid; c < G.n; c += kBatchThreads)
            {
                Tv dc;
                const Tv ax = mixRowAx<Tm, Tv>(G, c, Xg, &dc);
                const Tv xo = Xg(c) + Tv(kOmega) * (B(c) - ax) / safeDiag(dc);
                Yg.set(c, xo);
                if (last) inject(g, c, xo);
            }
            __syncthreads();
            flip ^= 1u << g;
        };
        for (int g = 0; g + 1 < T; ++g)
        {
            const MixGrid& G = M[g];
            for (int s = 1; s < J.nPre; ++s) sweep(g, false);
            const CycVec<Tv, Ts> Xg = vec(g, X(g)), B = rhs(G.b), Rg = rhs(G.r);
            if (J.nPre == 0)
            {
                for (int c = tid; c < G.n; c += kBatchThreads) Xg.set(c, Tv(0));
                __syncthreads();
            }
            for (int c = tid; c < G.n; c += kBatchThreads)
            {
                Tv dc;
                Rg.set(c, B(c) - mixRowAx<Tm, Tv>(G, c, Xg, &dc));
            }
            __syncthreads();
            const MixGrid& C = M[g + 1];
            const Tv* dC = static_cast<const Tv*>(C.diag);
            const CycVec<Tv, Ts> BC = rhs(C.b), XC = vec(g + 1, X(g + 1));
            const int fz = (g + 2 < T && J.nPre > 0) ? 1 : 0;
            for (int k = tid; k < C.n; k += kBatchThreads)
            {
                Tv s = Tv(0);
                for (int q = G.start(k); q < G.start(k + 1); ++q) s += Rg(G.list(q));
                BC.set(k, s);
                if (fz) XC.set(k, Tv(0) + Tv(kOmega) * (s - Tv(0)) / safeDiag(dC[k]));
            }
            __syncthreads();
        }
        {
            const MixGrid& G = M[T - 1];
            const CycVec<Tv, Ts> B = rhs(G.b), Xg = vec(T - 1, X(T - 1));
            for (int j = tid; j < G.n; j += kBatchThreads) sb[j] = (double)B(j);
            __syncthreads();
            for (int i = tid; i < G.n; i += kBatchThreads)
            {
                const double s = denseRowDot(J.inv + (size_t)i * G.n, sb, G.n);
                Xg.set(i, (Tv)s);
                inject(T - 1, i, (Tv)s);
            }
            __syncthreads();
        }
        const int nExtra = J.nPost > 1 ? J.nPost - 1 : 0;
        for (int g = T - 2; g >= 0; --g)
        {
            const MixGrid& G = M[g];
            const CycVec<Tv, Ts> E = vec(g, G.e);
            if (J.scaled)
            {
                // the scaled correction (tailK's): alpha = (r.e)/(e.Ae) over the workgroup, fused with the
                // first post-sweep, in place
                const CycVec<Tv, Ts> Xg = vec(g, X(g)), AE = rhs(Y(g)), Rg = rhs(G.r);
                double num = 0.0, den = 0.0;
                for (int c = tid; c < G.n; c += kBatchThreads)
                {
                    Tv dc;
                    const Tv ae = mixRowAx<Tm, Tv>(G, c, E, &dc);
                    AE.set(c, ae);
                    const double ev = (double)E(c);
                    num += ev * (double)Rg(c);
                    den += ev * (double)ae;
                }
                num = blockSum<kBatchThreads>(num, lds);
                den = blockSum<kBatchThreads>(den, lds);
                if (tid == 0) shA = den > 1e-30 ? num / den : 1.0;
                __syncthreads();
                const Tv al = (Tv)shA;
                const Tv* dg = static_cast<const Tv*>(G.diag);
                for (int c = tid; c < G.n; c += kBatchThreads)
                {
                    const Tv ev = E(c);
                    const Tv xo = J.nPost > 0 ? Xg(c) + al * ev + Tv(kOmega) * (Rg(c) - al * AE(c)) / safeDiag(dg[c])
                                              : Xg(c) + al * ev;
                    Xg.set(c, xo);
                    if (!nExtra) inject(g, c, xo);
                }
                __syncthreads();
                for (int s = 0; s < nExtra; ++s) sweep(g, s + 1 == nExtra);
            }
            else if (J.nPost > 0)
            {
                const CycVec<Tv, Ts> B = rhs(G.b), Yg = vec(g, Y(g));
                const SumX<Tv, Ts> XE{vec(g, X(g)), E};
                for (int c = tid; c < G.n; c += kBatchThreads)
                {
                    Tv dc;
                    const Tv ax = mixRowAx<Tm, Tv>(G, c, XE, &dc);
                    const Tv xo = XE(c) + Tv(kOmega) * (B(c) - ax) / safeDiag(dc);
                    Yg.set(c, xo);
                    if (!nExtra) inject(g, c, xo);
                }
                __syncthreads();
                flip ^= 1u << g;
                for (int s = 0; s < nExtra; ++s) sweep(g, s + 1 == nExtra);
            }
            else
            {
                const CycVec<Tv, Ts> Xg = vec(g, X(g));
                for (int c = tid; c < G.n; c += kBatchThreads)
                {
                    const Tv xo = Xg(c) + E(c);
                    Xg.set(c, xo);
                    inject(g, c, xo);
                }
                __syncthreads();
            }
        }
        // ---- z.r (and z.r_old), the Polak-Ribiere+ beta ----
        const CycVec<Tv, Ts> Z = vec(0, X(0));
        auto z = [&](int c) { return sig != 0.0 ? (double)Z(c) * sig : (double)Z(c); };
        double zr = 0.0, zo = 0.0;
        for (int c = tid; c < n; c += kBatchThreads)
        {
            const double zc = z(c);
            zr += zc * r[c];
            if (J.flexible && it > 0) zo += zc * rOld[c];
        }
        zr = blockSum<kBatchThreads>(zr, lds);
        zo = blockSum<kBatchThreads>(zo, lds);
        if (tid == 0) shA = zr - zo;
        if (tid == 0) shB = zr;
        __syncthreads();
        const double rho = shB;
        double beta = 0.0;
        if (it > 0)
        {
            if (J.flexible)
                beta = (rhoOld > 1e-300 || rhoOld < -1e-300) ? fmax(0.0, shA / rhoOld) : 0.0;
            else
                beta = fabs(rhoOld) > 1e-300 ? rho / rhoOld : 0.0;
        }
        rhoOld = rho;
        // ---- p = z + beta p, q = A p, p.q ----
        for (int c = tid; c < n; c += kBatchThreads)
        {
            const double zc = z(c);
            J.p[c] = it > 0 ? beta * J.p[c] + zc : zc;
        }
        __syncthreads();
        double pq = 0.0;
        for (int c = tid; c < n; c += kBatchThreads)
        {
            const double pc = J.p[c];
            const double ax = rowOffSum<double, double>(J.S0, c, A, PlainX<double>{J.p}, Dg[c] * pc);
            J.q[c] = ax;
            pq += ax * pc;
        }
        pq = blockSum<kBatchThreads>(pq, lds);
        if (tid == 0) shA = pq;
        __syncthreads();
        const double alpha = fabs(shA) > 1e-300 ? rho / shA : 0.0;
        // ---- x += alpha p, r -= alpha q (into the other buffer), sum|r| ----
        double* rn = rPrev;   // the next residual overwrites the one before this
        double sr = 0.0;
        for (int c = tid; c < n; c += kBatchThreads)
        {
            x[c] += alpha * J.p[c];
            const double v = r[c] - alpha * J.q[c];
            rn[c] = v;
            sr += fabs(v);
        }
        sr = blockSum<kBatchThreads>(sr, lds);
        if (tid == 0) shA = sr;
        __syncthreads();
        res = shA / nf;
        rPrev = rCur;
        rCur = rn;
        ++it;
        conv = converged(res);
        stop = ((conv || it >= J.maxIter) && it >= J.minIter);
    }
    if (tid == 0)
    {
        J.outD[0] = init;
        J.outD[1] = res;
        J.outI[0] = it;
        J.outI[1] = conv ? 1 : 0;
    }
}

// ---------------- Galerkin and fill kernels ----------------

// coarse diag = sum(fine diag, ascending fine cell) + sum(upper + lower of the fine faces inside the
// agglomerate, ascending) (galDiagGatherK)
__global__ __launch_bounds__(kTpb) void galDiagK(int nc, const int* __restrict__ cellStart,
                                                 const int* __restrict__ cellList,
                                                 const int* __restrict__ dfaceStart,
                                                 const int* __restrict__ dfaceList, const double* __restrict__ fd,
                                                 const double* __restrict__ fu, double* __restrict__ cd)
{
    const int C = xcdBlock() * kTpb + threadIdx.x;
    if (C >= nc) return;
    double s = 0.0;
    for (int k = cellStart[C]; k < cellStart[C + 1]; ++k) s += fd[cellList[k]];
    for (int k = dfaceStart[C]; k < dfaceStart[C + 1]; ++k)
    {
        const int f = dfaceList[k];
        s += fu[f] + fu[f];
    }
    cd[C] = s;
}

// coarse face = sum of its fine faces, ascending (galFaceGatherSymK)
__global__ __launch_bounds__(kTpb) void galFaceK(int ncf, const int* __restrict__ faceStart,
                                                 const int* __restrict__ faceList, const double* __restrict__ fu,
                                                 double* __restrict__ cu)
{
    const int F = xcdBlock() * kTpb + threadIdx.x;
    if (F >= ncf) return;
    double u = 0.0;
    for (int k = faceStart[F]; k < faceStart[F + 1]; ++k) u += fu[faceList[k]];
    cu[F] = u;
}

// sliced-row copy of a grid's face values in precision Tm, scaled by 1/s (s from the device)
template <typename Tm>
__global__ __launch_bounds__(kTpb) void fillK(int nf, const int* __restrict__ slotLo, const int* __restrict__ slotHi,
                                              const double* __restrict__ up, const double* __restrict__ lo,
                                              const double* __restrict__ scale, Tm* __restrict__ val)
{
    const int f = xcdBlock() * kTpb + threadIdx.x;
    if (f >= nf) return;
    const double inv = scale ? 1.0 / *scale : 1.0;
    if constexpr (sizeof(Tm) == 2)
    {
        val[slotLo[f]] = __float2half((float)(up[f] * inv));
        val[slotHi[f]] = __float2half((float)(lo[f] * inv));
    }
    else
    {
        val[slotLo[f]] = (Tm)up[f];
        val[slotHi[f]] = (Tm)lo[f];
    }
}

template <typename Tv>
__global__ __launch_bounds__(kTpb) void castK(int n, const double* __restrict__ a, Tv* __restrict__ b)
{
    const int i = xcdBlock() * kTpb + threadIdx.x;
    if (i < n) b[i] = (Tv)a[i];
}

// grid 0's diagonal copy: each face's upper coefficient into its owner's slot of its offset, in
// precision Tm, scaled by 1/s (s from the device) as fillK
template <typename Tm>
__global__ __launch_bounds__(kTpb) void diaFillK(int nf, const int* __restrict__ slot, const double* __restrict__ up,
                                                 const double* __restrict__ scale, Tm* __restrict__ dia)
{
    const int f = xcdBlock() * kTpb + threadIdx.x;
    if (f >= nf) return;
    if constexpr (sizeof(Tm) == 2)
        dia[slot[f]] = __float2half((float)(up[f] * (1.0 / *scale)));
    else
        dia[slot[f]] = (Tm)up[f];
}

// the cycle's grid-0 result in FP64 (the FP16 cycle's times its input scale)
template <typename Tv, typename Ts>
__global__ __launch_bounds__(kTpb) void widenK(int n, const Ts* __restrict__ a, const float* __restrict__ xm,
                                               const double* __restrict__ sig, double* __restrict__ b)
{
    const int i = xcdBlock() * kTpb + threadIdx.x;
    if (i >= n) return;
    const CycVec<Tv, Ts> A(a, xm);
    b[i] = sig ? (double)A(i) * *sig : (double)A(i);
}

// per-grid FP16 scale: the smallest power of two >= max|upper| (exact scaling); max is exact, so the
// combine order does not matter
__global__ __launch_bounds__(kReduceBlock) void maxAbsK(int n, const double* __restrict__ a,
                                                       double* __restrict__ partials)
{
    __shared__ double lds[kReduceBlock / 64];
    double v = 0.0;
    MFV_ROWS(n, i) v = fmax(v, fabs(a[i]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double m = 0.0;
        for (int w = 0; w < kReduceBlock / 64; ++w) m = fmax(m, lds[w]);
        partials[blockIdx.x] = m;
    }
}

__global__ __launch_bounds__(kFinishBlock) void maxPow2FinishK(int nParts, const double* __restrict__ partials,
                                                                double* __restrict__ out)
{
    __shared__ double lds[kFinishBlock / 64];
    double mm = finishMax(nParts, partials, 0.0);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mm = fmax(mm, __shfl_xor(mm, o, 64));
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = mm;
    __syncthreads();
    if (threadIdx.x) return;
    for (int w = 0; w < kFinishBlock / 64; ++w) mm = fmax(mm, lds[w]);
    *out = mm > 0.0 ? exp2(ceil(log2(mm))) : 1.0;
}

// the FP16 cycle's input scale from max|r| (partials of maxAbsK): sig[0] (this cycle) = sig[1] (the next
// one's, until the solver's update replaces it) = the power of two at or above the max
__global__ __launch_bounds__(kFinishBlock) void sigFinishK(int nParts, const double* __restrict__ partials,
                                                           double* __restrict__ sig)
{
    __shared__ double lds[kFinishBlock / 64];
    double mm = blockMax<kFinishBlock>(finishMax(nParts, partials, 0.0), lds);
    if (threadIdx.x) return;
    const double s = mm > 0.0 ? exp2(ceil(log2(mm))) : 1.0;
    sig[0] = s;
    sig[1] = s;
}

// Gauss-Jordan inverse with partial pivoting (first largest |a_ik|, i >= k) on [A | I], the
// reference's pivot floor 1e-14*max|diag|. One workgroup; every element update is the sequential
// algorithm's own, so the result is bit-identical to a one-thread loop.
__global__ __launch_bounds__(kTpb) void coarseInverseK(int n, int nf, const int* __restrict__ own,
                                                       const int* __restrict__ nei, const double* __restrict__ diag,
                                                       const double* __restrict__ upper, double* __restrict__ M,
                                                       double* __restrict__ inv)
{
    const int tid = threadIdx.x;
    const int w = 2 * n;
    __shared__ double fcol[kMaxDenseCoarsest];
    __shared__ double shScale, shPiv;
    __shared__ int shRow;
    for (int i = tid; i < n * w; i += kTpb) M[i] = 0.0;
    __syncthreads();
    for (int c = tid; c < n; c += kTpb)
    {
        M[c * w + c] = diag[c];
        M[c * w + n + c] = 1.0;
    }
    __syncthreads();
    if (tid == 0)
    {
        for (int f = 0; f < nf; ++f)
        {
            M[own[f] * w + nei[f]] += upper[f];
            M[nei[f] * w + own[f]] += upper[f];
        }
        double m = 0.0;
        for (int c = 0; c < n; ++c) m = fmax(m, fabs(diag[c]));
        shScale = m > 0.0 ? m : 1.0;
    }
    __syncthreads();
    for (int k = 0; k < n; ++k)
    {
        if (tid == 0)
        {
            double bv = -1.0;
            int bi = k;
            for (int i = k; i < n; ++i)
            {
                const double v = fabs(M[i * w + k]);
                if (v > bv)
                {
                    bv = v;
                    bi = i;
                }
            }
            shRow = bi;
        }
        __syncthreads();
        const int p = shRow;
        if (p != k)
            for (int j = tid; j < w; j += kTpb)
            {
                const double t = M[k * w + j];
                M[k * w + j] = M[p * w + j];
                M[p * w + j] = t;
            }
        __syncthreads();
        if (tid == 0)
        {
            double d = M[k * w + k];
            const double fl = 1e-14 * shScale;
            if (!(fabs(d) > fl)) d = d < 0.0 ? -fl : fl;
            M[k * w + k] = d;
            shPiv = d;
        }
        __syncthreads();
        const double rp = 1.0 / shPiv;
        for (int j = tid; j < w; j += kTpb) M[k * w + j] *= rp;
        __syncthreads();
        for (int i = tid; i < n; i += kTpb) fcol[i] = M[i * w + k];
        __syncthreads();
        for (int idx = tid; idx < n * w; idx += kTpb)
        {
            const int i = idx / w, j = idx - i * w;
            if (i == k) continue;
            M[i * w + j] -= fcol[i] * M[k * w + j];
        }
        __syncthreads();
    }
    for (int idx = tid; idx < n * n; idx += kTpb)
    {
        const int i = idx / n, j = idx - i * n;
        inv[idx] = M[i * w + n + j];
    }
}

// y = A x in FP64 on the sliced rows (the fine grid's FP64 copy)
__global__ __launch_bounds__(kTpb) void sellAmul64K(SellView S, const double* __restrict__ val,
                                                    const double* __restrict__ diag, const double* __restrict__ x,
                                                    double* __restrict__ y)
{
    const int c = xcdBlock() * kTpb + threadIdx.x;
    if (c >= S.n) return;
    y[c] = rowOffSum<double, double>(S, c, val, PlainX<double>{x}, diag[c] * x[c]);
}

// ---------------- host side ----------------

void buildSell(Ctx& c, Sell& S, int n, const std::vector<int>& own, const std::vector<int>& nei)
{
    const int nF = (int)own.size();
    std::vector<int> os(n + 1, 0), ls(n + 1, 0), lo(nF);
    for (int f = 0; f < nF; ++f)
    {
        ++os[own[f] + 1];
        ++ls[nei[f] + 1];
    }
    for (int i = 0; i < n; ++i)
    {
        os[i + 1] += os[i];
        ls[i + 1] += ls[i];
    }
    {
        std::vector<int> at(ls.begin(), ls.end() - 1);
        for (int f = 0; f < nF; ++f) lo[at[nei[f]]++] = f;
    }
    const int nSl = (n + kSlice - 1) / kSlice;
    std::vector<int> slicePtr(nSl + 1, 0);
    int width = 0;
    for (int s = 0; s < nSl; ++s)
    {
        int wmax = 0;
        for (int r = s * kSlice; r < std::min(n, (s + 1) * kSlice); ++r)
            wmax = std::max(wmax, (ls[r + 1] - ls[r]) + (os[r + 1] - os[r]));
        width = std::max(width, wmax);
        const long long next = (long long)slicePtr[s] + (long long)wmax * kSlice;
        if (next > 0x7fffffffLL) invalid("sliced rows exceed 2^31 entries");
        slicePtr[s + 1] = (int)next;
    }
    const long long nnz = slicePtr[nSl];
    std::vector<int> cols(nnz), slotLo(nF), slotHi(nF);
    std::vector<uint8_t> pat(n, 0);
    std::map<std::vector<int>, int> patterns;
    bool patterned = true;
    std::vector<int> offs, lastOffs;
    int lastPat = -1;
    for (int r = 0; r < n; ++r)
    {
        const int s = r / kSlice, lane = r % kSlice;
        const int w = (slicePtr[s + 1] - slicePtr[s]) / kSlice;
        int k = 0;
        offs.clear();
        for (int q = ls[r]; q < ls[r + 1]; ++q, ++k)
        {
            const int f = lo[q];
            const int e = slicePtr[s] + k * kSlice + lane;
            cols[e] = own[f];
            slotHi[f] = e;
            offs.push_back(own[f] - r);
        }
        for (int f = os[r]; f < os[r + 1]; ++f, ++k)
        {
            const int e = slicePtr[s] + k * kSlice + lane;
            cols[e] = nei[f];
            slotLo[f] = e;
            offs.push_back(nei[f] - r);
        }
        for (; k < w; ++k) cols[slicePtr[s] + k * kSlice + lane] = r;
        if (patterned)
        {
            // neighbouring rows mostly share a pattern: the previous row's is tried before the map
            if (lastPat >= 0 && offs == lastOffs)
                pat[r] = (uint8_t)lastPat;
            else
            {
                auto it = patterns.find(offs);
                if (it == patterns.end())
                {
                    if (patterns.size() >= 255)
                        patterned = false;
                    else
                        it = patterns.emplace(offs, (int)patterns.size()).first;
                }
                if (patterned)
                {
                    pat[r] = (uint8_t)it->second;
                    lastPat = it->second;
                    lastOffs = offs;
                }
            }
        }
    }
    // rows of the last slice beyond n: padding points at row 0 (never read: threads exit)
    for (int r = n; r < nSl * kSlice; ++r)
    {
        const int s = r / kSlice, lane = r % kSlice;
        const int w = (slicePtr[s + 1] - slicePtr[s]) / kSlice;
        for (int k = 0; k < w; ++k) cols[slicePtr[s] + k * kSlice + lane] = 0;
    }
    S.n = n;
    S.nSlices = nSl;
    S.width = width;
    S.nnz = nnz;
    S.patterned = patterned && width > 0;
    S.slicePtr.upload(slicePtr, c.stream);
    if (S.patterned)
    {
        std::vector<int> table(patterns.size() * (size_t)width, 0);
        for (const auto& kv : patterns)
            for (size_t k = 0; k < kv.first.size(); ++k) table[(size_t)kv.second * width + k] = kv.first[k];
        S.pattern.upload(pat, c.stream);
        S.patOff.upload(table, c.stream);
    }
    else
    {
        S.cols.upload(cols, c.stream);
    }
    S.slotLo.upload(slotLo, c.stream);
    S.slotHi.upload(slotHi, c.stream);
    hipCheck(hipStreamSynchronize(c.stream), "sell upload");
}

template <typename T>
void zeroBuf(Ctx& c, DBuf<T>& b, size_t n)
{
    b.resize(n);
    if (n) hipCheck(hipMemsetAsync(b.p, 0, n * sizeof(T), c.stream), "zero");
}

// Grid 0's diagonal copy, when its faces form a fixed stencil of at most kDiaMax positive offsets in
// (owner, neighbour) order and the diagonals hold at most two slots per face (a hex mesh: three
// offsets, about one slot per face). Otherwise diaM stays 0 and the grid keeps its sliced rows only.
void buildDia(Ctx& c, Grid& g, const std::vector<int>& own, const std::vector<int>& nei)
{
    g.diaM = 0;
    const int nF = (int)own.size();
    if (nF == 0) return;
    std::vector<int> offs;
    for (int f = 0; f < nF; ++f)
    {
        if (f > 0 && (own[f] < own[f - 1] || (own[f] == own[f - 1] && nei[f] <= nei[f - 1]))) return;
        const int d = nei[f] - own[f];
        if (std::find(offs.begin(), offs.end(), d) == offs.end())
        {
            if ((int)offs.size() == kDiaMax) return;
            offs.push_back(d);
        }
    }
    std::sort(offs.begin(), offs.end());
    const long long pad = ((long long)g.n + kSlice - 1) / kSlice * kSlice;
    const long long slots = (long long)offs.size() * pad;
    if (slots > 2LL * nF + (long long)kDiaMax * kSlice || slots > 0x7fffffffLL) return;
    std::vector<int> slot(nF);
    for (int f = 0; f < nF; ++f)
    {
        const int k = (int)(std::lower_bound(offs.begin(), offs.end(), nei[f] - own[f]) - offs.begin());
        slot[f] = (int)(k * pad + own[f]);
    }
    g.diaM = (int)offs.size();
    for (int k = 0; k < g.diaM; ++k) g.diaOff[k] = offs[k];
    g.diaPad = (int)pad;
    g.diaSlot.upload(slot, c.stream);
    zeroBuf(c, g.dia64, (size_t)slots);
}

void restrictionLists(Ctx& c, Grid& G, const AggLevel& L)
{
    const int nC = L.nCoarse, nCF = L.nCoarseFaces, nF = L.nFineFaces;
    std::vector<int> cs(nC + 1, 0), cl(L.nFine), ds(nC + 1, 0), fs(nCF + 1, 0);
    for (int i = 0; i < L.nFine; ++i) ++cs[L.map[i] + 1];
    for (int i = 0; i < nC; ++i) cs[i + 1] += cs[i];
    {
        std::vector<int> at(cs.begin(), cs.end() - 1);
        for (int i = 0; i < L.nFine; ++i) cl[at[L.map[i]]++] = i;
    }
    for (int f = 0; f < nF; ++f)
    {
        const int r = L.faceRestrict[f];
        if (r >= 0)
            ++fs[r + 1];
        else
            ++ds[-1 - r + 1];
    }
    for (int i = 0; i < nC; ++i) ds[i + 1] += ds[i];
    for (int i = 0; i < nCF; ++i) fs[i + 1] += fs[i];
    std::vector<int> dl(ds[nC]), fl(fs[nCF]);
    {
        std::vector<int> dat(ds.begin(), ds.end() - 1), fat(fs.begin(), fs.end() - 1);
        for (int f = 0; f < nF; ++f)
        {
            const int r = L.faceRestrict[f];
            if (r >= 0)
                fl[fat[r]++] = f;
            else
                dl[dat[-1 - r]++] = f;
        }
    }
    std::vector<int4> qd(nC);
    for (int C = 0; C < nC; ++C)
    {
        const int k = cs[C + 1] - cs[C];
        int v[4] = {-1, -1, -1, -1};
        if (k <= 4)
            for (int j = 0; j < k; ++j) v[j] = cl[cs[C] + j];
        else
        {
            v[0] = -1;
            v[1] = C;
        }
        qd[C] = int4{v[0], v[1], v[2], v[3]};
    }
    G.map.upload(L.map, c.stream);
    G.cellStart.upload(cs, c.stream);
    G.cellList.upload(cl, c.stream);
    G.quads.upload(qd, c.stream);
    G.dfaceStart.upload(ds, c.stream);
    G.dfaceList.upload(dl, c.stream);
    G.faceStart.upload(fs, c.stream);
    G.faceList.upload(fl, c.stream);
}

// grid 0's V-cycle reads its diagonal copies when the stencil has two or three offsets (a 2-D or 3-D
// structured mesh); any other grid, or stencil, its sliced rows
bool cycleDia(const Grid& g) { return g.diaM == 2 || g.diaM == 3; }

SellView cycleView(const Grid& g, bool dia)
{
    SellView S = viewOf(g.sell);
    if (dia)
    {
        S.diaM = g.diaM;
        S.diaPad = g.diaPad;
        for (int k = 0; k < kDiaMax; ++k) S.diaOff[k] = g.diaOff[k];
    }
    return S;
}

template <typename Tm>
DBuf<Tm>& diaOf(Grid& g);
template <>
DBuf<double>& diaOf<double>(Grid& g) { return g.dia64; }
template <>
DBuf<float>& diaOf<float>(Grid& g) { return g.dia32; }
template <>
DBuf<__half>& diaOf<__half>(Grid& g) { return g.dia16; }

template <typename Tm>
DBuf<Tm>& offOf(Grid& g);
template <>
DBuf<double>& offOf<double>(Grid& g) { return g.off64; }
template <>
DBuf<float>& offOf<float>(Grid& g) { return g.off32; }
template <>
DBuf<__half>& offOf<__half>(Grid& g) { return g.off16; }

template <typename Tv, typename Ts>
struct VecSet
{
    Ts *b, *x, *y, *r, *e;
    const Tv* diag;
    const float* xm;   // solution-vector multiplier (FP16 cycle), else null
};

VecSet<double, double> vecs64(Amg&, Grid& g)
{
    return {g.b64.p, g.x64.p, g.y64.p, g.r64.p, g.e64.p, g.diagV64.p, nullptr};
}
VecSet<float, float> vecs32(Amg&, Grid& g)
{
    return {g.b32.p, g.x32.p, g.y32.p, g.r32.p, g.e32.p, g.diagV32.p, nullptr};
}
VecSet<float, __half> vecs16(Amg& a, Grid& g)
{
    return {g.b16.p, g.x16.p, g.y16.p, g.r16.p, g.e16.p, g.diagV32.p, a.xmul.p + (&g - a.grids.data())};
}

// Runs the cycle from grid gi down and back. Grid gi's final values go to up (the finer grid's e, or
// nowhere for grid 0); the return value is the buffer that holds them (x or y: the ping-pong sweeps
// leave it in either).
template <typename Tm, typename Tv, typename Ts>
Ts* vcycleAt(Ctx& c, Amg& a, int gi, const VcycleControls& vc, const int* stop,
             VecSet<Tv, Ts> (*vs)(Amg&, Grid&), Scatter<Tv, Ts> up)
{
    Grid& g = a.grids[gi];
    VecSet<Tv, Ts> v = vs(a, g);
    const int n = g.n;
    const int last = (int)a.grids.size() - 1;
    const Scatter<Tv, Ts> none{nullptr, nullptr, nullptr, nullptr, nullptr};
    if (gi == last)
    {
        ProfScope ps(c, kProfGridDown + gi);
        coarseApplyK<Tv, Ts><<<1, kTpb, 0, c.stream>>>(n, a.coarseInv.p, v.b, v.x, v.xm, up, stop);
        return v.x;
    }
    if (gi > 0 && gi == a.tailStart && !c.sequentialReduce && a.coarseMode > 0)
    {
        ProfScope ps(c, kProfGridDown + gi);
        if (a.coarseMode == 2 && a.tailLdsBytes > 0 && !vc.scale)
        {
            tailMixK<Tm, Tv, Ts><<<1, kTailThreads, a.tailLdsBytes, c.stream>>>(
                reinterpret_cast<const TailGrid*>(a.tailDesc.p), reinterpret_cast<const LdsTail*>(a.tailLdsDesc.p),
                a.coarseInv.p, vc.nPre, vc.nPost, up, stop);
        }
        else
            tailK<Tm, Tv, Ts><<<1, kTailThreads, 0, c.stream>>>(reinterpret_cast<const TailGrid*>(a.tailDesc.p),
                                                                 last - gi + 1, a.coarseInv.p, vc.nPre, vc.nPost,
                                                                 vc.scale ? 1 : 0, up, stop);
        hipCheck(hipGetLastError(), "vcycle tail");
        return nullptr;   // grid gi's values went up through `up`; no caller reads them here
    }
    ProfScope psDown(c, kProfGridDown + gi);
    const bool dia = gi == 0 && cycleDia(g);
    const SellView S = cycleView(g, dia);
    const Tm* val = dia ? diaOf<Tm>(g).p : offOf<Tm>(g).p;
    const double* sc = a.scales.p + gi;
    // x -> y by one Jacobi sweep
    auto smooth = [&](const Ts* x, Ts* y, const Scatter<Tv, Ts>& to) {
        smoothK<Tm, Tv, Ts><<<blocks(n), kTpb, 0, c.stream>>>(S, val, v.diag, sc, x, v.b, y, v.xm, to, stop);
    };
    // pre-sweeps after the first (the first ran from zero where b was made)
    for (int s = 1; s < vc.nPre; ++s)
    {
        smooth(v.x, v.y, none);
        std::swap(v.x, v.y);
    }
    if (vc.nPre == 0) zeroK<Ts><<<blocks(n), kTpb, 0, c.stream>>>(n, v.x, stop);
    {
        ProfScope ps(c, gi == 0 ? kProfAmgResidual0 : kProfCount);
        residualK<Tm, Tv, Ts><<<blocks(n), kTpb, 0, c.stream>>>(S, val, v.diag, sc, v.x, v.b, v.r, v.xm, stop);
    }
    Grid& gc = a.grids[gi + 1];
    VecSet<Tv, Ts> w = vs(a, gc);
    const int coarseIsLast = (gi + 1 == last);
    {
        ProfScope ps(c, gi == 0 ? kProfAmgRestrict0 : kProfCount);
        if (!coarseIsLast && vc.nPre > 0)
            restrictK<Tv, Ts, true><<<blocks(gc.n), kTpb, 0, c.stream>>>(gc.n, g.cellStart.p, g.cellList.p, g.quads.p,
                                                                          v.r, w.b, w.diag, w.x, w.xm, stop);
        else
            restrictK<Tv, Ts, false><<<blocks(gc.n), kTpb, 0, c.stream>>>(gc.n, g.cellStart.p, g.cellList.p,
                                                                           g.quads.p, v.r, w.b, w.diag, w.x, w.xm, stop);
    }
    psDown.close();
    // the coarse grid leaves its result, injected, in this grid's e
    vcycleAt<Tm, Tv, Ts>(c, a, gi + 1, vc, stop, vs,
                         Scatter<Tv, Ts>{g.cellStart.p, g.cellList.p, v.e, v.xm, g.quads.p});
    ProfScope psUp(c, kProfGridUp + gi);
    // the kernel that finishes this grid hands its values on to the finer grid
    const int nExtra = std::max(0, vc.nPost - 1);
    if (vc.scale)
    {
        const bool seq = c.sequentialReduce;
        // four rows per thread on the big grids (measured faster at 20 M rows), two below
        const int gr = reduceGridRows(n, n >= (1 << 22) ? 4 : 2);
        double* p1 = c.partials.p + 1 * (size_t)kMaxReduceBlocks;
        double* p2 = c.partials.p + 2 * (size_t)kMaxReduceBlocks;
        {
            ProfScope ps(c, gi == 0 ? kProfAmgCorrAmul0 : kProfCount);
            corrAmulK<Tm, Tv, Ts><<<gr, kReduceBlock, 0, c.stream>>>(S, val, v.diag, sc, v.e, v.r, v.y, v.xm, p1, p2,
                                                                     seq ? a.seqA.p : nullptr,
                                                                     seq ? a.seqB.p : nullptr, stop);
        }
        const Scatter<Tv, Ts> to = nExtra ? none : up;
        // four rows a thread pays where nothing is scattered (grid 0); a grid that injects its values
        // into the finer one keeps a thread per row for the scattered writes
        bool wide = false;
        if constexpr (std::is_same<Ts, __half>::value) wide = to.e == nullptr;
        // a small grid's alpha is finished by every workgroup of the update itself (same bits, one launch
        // fewer on a latency-bound grid)
        const bool alphaInBlock = !seq && !wide && gr <= kAlphaInBlockParts;
        if (seq)
            seqFinish2K<<<1, 64, 0, c.stream>>>(n, a.seqA.p, a.seqB.p, AlphaPost{a.alphas.p + gi}, stop);
        else if (!alphaInBlock)
            finish2K<<<1, kFinishBlock, 0, c.stream>>>(gr, p1, p2, AlphaPost{a.alphas.p + gi}, stop);
        {
            ProfScope ps(c, gi == 0 ? kProfAmgCorrSmooth0 : kProfCount);
            if constexpr (std::is_same<Ts, __half>::value)
            {
                if (wide)
                    corrSmoothH4K<<<blocks((n + 3) / 4), kTpb, 0, c.stream>>>(n, v.e, v.r, v.y, v.diag, a.alphas.p + gi,
                                                                               vc.nPost > 0 ? 1 : 0, v.x, v.xm, to,
                                                                               stop);
            }
            if (!wide)
                corrSmoothK<Tv, Ts><<<blocks(n), kTpb, 0, c.stream>>>(n, v.e, v.r, v.y, v.diag, a.alphas.p + gi,
                                                                      vc.nPost > 0 ? 1 : 0, v.x, v.xm, to, stop,
                                                                      alphaInBlock ? gr : 0, p1, p2);
        }
        for (int s = 0; s < nExtra; ++s)
        {
            smooth(v.x, v.y, s + 1 == nExtra ? up : none);
            std::swap(v.x, v.y);
        }
    }
    else if (vc.nPost > 0)
    {
        {
            ProfScope ps(c, gi == 0 ? kProfAmgCorrSmooth0 : kProfCount);
            prolongSmoothK<Tm, Tv, Ts><<<blocks(n), kTpb, 0, c.stream>>>(S, val, v.diag, sc, v.e, v.x, v.b, v.y, v.xm,
                                                                          nExtra ? none : up, stop);
        }
        std::swap(v.x, v.y);
        for (int s = 0; s < nExtra; ++s)
        {
            smoothK<Tm, Tv, Ts><<<blocks(n), kTpb, 0, c.stream>>>(S, val, v.diag, sc, v.x, v.b, v.y, v.xm,
                                                                  s + 1 == nExtra ? up : none, stop);
            std::swap(v.x, v.y);
        }
    }
    else
    {
        prolongK<Tv, Ts><<<blocks(n), kTpb, 0, c.stream>>>(n, v.e, v.x, v.xm, up, stop);
    }
    hipCheck(hipGetLastError(), "vcycle");
    return v.x;
}

template <typename Tv>
void allocVectors(Ctx& c, Grid& g);
template <>
void allocVectors<double>(Ctx& c, Grid& g)
{
    zeroBuf(c, g.b64, g.n);
    zeroBuf(c, g.x64, g.n);
    zeroBuf(c, g.y64, g.n);
    zeroBuf(c, g.r64, g.n);
    zeroBuf(c, g.e64, g.n);
}
template <>
void allocVectors<float>(Ctx& c, Grid& g)
{
    zeroBuf(c, g.b32, g.n);
    zeroBuf(c, g.x32, g.n);
    zeroBuf(c, g.y32, g.n);
    zeroBuf(c, g.r32, g.n);
    zeroBuf(c, g.e32, g.n);
}
template <>
void allocVectors<__half>(Ctx& c, Grid& g)
{
    zeroBuf(c, g.b16, g.n);
    zeroBuf(c, g.x16, g.n);
    zeroBuf(c, g.y16, g.n);
    zeroBuf(c, g.r16, g.n);
    zeroBuf(c, g.e16, g.n);
}

} // namespace

void amgBuild(Ctx& c, const Mesh& m, Amg& a, const std::vector<double>& faceWeights, const AmgBuildControls& ctl)
{
    if ((int)faceWeights.size() != m.nFaces) invalid("amg build: one face weight per internal face is required");
    if (ctl.target < 1) invalid("amg build: the coarsest-level target must be positive");
    bool forward = true;
    amgBuildFrom(c, m, a,
                 mergeLevels(agglomerate(ctl.kind, m.nCells, m.hOwner, m.hNeighbour, faceWeights, ctl.target,
                                         forward),
                             ctl.mergeLevels),
                 ctl);
}

void amgBuildFrom(Ctx& c, const Mesh& m, Amg& a, std::vector<AggLevel> levels, const AmgBuildControls& ctl)
{
    a.nFine = m.nCells;
    a.controls = ctl;
    a.agg = std::move(levels);
    for (size_t i = 0; i < a.agg.size(); ++i)
    {
        const AggLevel& L = a.agg[i];
        const int nF = i == 0 ? m.nCells : a.agg[i - 1].nCoarse;
        const int nFF = i == 0 ? m.nFaces : a.agg[i - 1].nCoarseFaces;
        if (L.nFine != nF || L.nFineFaces != nFF || (int)L.map.size() != nF || (int)L.faceRestrict.size() != nFF ||
            (int)L.faceFlip.size() != nFF || (int)L.cOwn.size() != L.nCoarseFaces ||
            (int)L.cNei.size() != L.nCoarseFaces)
            invalid("amg: level " + std::to_string(i) + " does not fit the grid above it");
        for (int v : L.map)
            if (v < 0 || v >= L.nCoarse) invalid("amg: level " + std::to_string(i) + " maps outside its coarse grid");
        for (int f = 0; f < L.nCoarseFaces; ++f)
            if (L.cOwn[f] < 0 || L.cOwn[f] >= L.cNei[f] || L.cNei[f] >= L.nCoarse)
                invalid("amg: level " + std::to_string(i) + " has a coarse face outside upper-triangular order");
        for (int r : L.faceRestrict)
            if (r >= L.nCoarseFaces || r < -L.nCoarse)
                invalid("amg: level " + std::to_string(i) + " restricts a face outside its coarse grid");
    }
    const int nG = (int)a.agg.size() + 1;
    a.grids.clear();
    a.grids.resize(nG);
    for (int gi = 0; gi < nG; ++gi)
    {
        Grid& g = a.grids[gi];
        const std::vector<int>& own = gi == 0 ? m.hOwner : a.agg[gi - 1].cOwn;
        const std::vector<int>& nei = gi == 0 ? m.hNeighbour : a.agg[gi - 1].cNei;
        g.n = gi == 0 ? m.nCells : a.agg[gi - 1].nCoarse;
        g.nFaces = (int)own.size();
        g.hOwn = own;
        g.hNei = nei;
        if (gi < nG - 1) restrictionLists(c, g, a.agg[gi]);
        if (gi > 0)
        {
            g.diag64.resize(g.n);
            g.upper64.resize(g.nFaces);
        }
        if (gi < nG - 1 || gi == 0) buildSell(c, g.sell, g.n, own, nei);
        if (gi == 0 && !ctl.fineRows) buildDia(c, g, own, nei);
    }
    a.nCoarsest = a.grids.back().n;
    Grid& gl = a.grids.back();
    gl.diag64.resize(gl.n);
    gl.upper64.resize(gl.nFaces);
    gl.denseOwn.upload(gl.hOwn, c.stream);
    gl.denseNei.upload(gl.hNei, c.stream);
    a.scales.resize(nG);
    std::vector<double> ones(nG, 1.0);
    a.scales.upload(ones, c.stream);
    a.alphas.resize(nG);
    a.seqA.resize(m.nCells);
    a.seqB.resize(m.nCells);
    a.galerkinDone = false;
    a.allocatedVectors = false;
    hipCheck(hipStreamSynchronize(c.stream), "amg build");
}

namespace {

// The tail's per-grid descriptors for the current precision (pointers into the grids' buffers, which
// stay put until the next allocation): grids from the first one of at most kTailCells cells (never
// grid 0) to the coarsest.
template <typename T>
std::vector<T> download(Ctx& c, const T* p, size_t n)
{
    std::vector<T> h(n);
    if (n) hipCheck(hipMemcpyAsync(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost, c.stream), "tail download");
    hipCheck(hipStreamSynchronize(c.stream), "tail download");
    return h;
}

// The LDS-staged part of the tail (tailMixK) for the current precision: the longest run of coarsest
// grids whose arrays fit the LDS a workgroup may hold beside the kernel's own, its static blob (slice
// starts, 16-bit columns -- a patterned grid's written out -- and 16-bit restriction lists) and its LDS
// layout. tailLdsBytes stays 0 when not even the two coarsest grids fit.
void buildTailLds(Ctx& c, Amg& a)
{
    a.tailLdsBytes = 0;
    a.tailLdsGrid = 0;
    const int nG = (int)a.grids.size();
    const int T = a.tailStart ? nG - a.tailStart : 0;
    if (T < 2) return;
    const bool f64 = a.prec == Prec::F64, f32 = a.prec == Prec::F32;
    const size_t sm = f64 ? 8 : f32 ? 4 : 2;   // coefficient
    const size_t sv = f64 ? 8 : 4;             // arithmetic (diagonal)
    const size_t ss = f64 ? 8 : f32 ? 4 : 2;   // stored vector
    auto al = [](size_t x) { return (x + 15) & ~(size_t)15; };
    auto u16 = [](int v) { return (unsigned short)(v >= 0 && v <= 65535 ? v : 0); };
    int maxShared = 0;
    hipCheck(hipDeviceGetAttribute(&maxShared, hipDeviceAttributeMaxSharedMemoryPerBlock, c.device), "lds size");
    const size_t own = sizeof(MixGrid) * kMaxTail + sizeof(StageSeg) * kMaxStageSegs + 256;   // the kernel's static LDS
    if ((size_t)maxShared <= own) return;
    const size_t budget = (size_t)maxShared - own;
    // per grid: the static arrays and the staged ones
    struct Host
    {
        std::vector<int> sp;
        std::vector<unsigned short> cols, cs, cl;
        size_t nnz = 0;
    };
    std::vector<Host> h(T);
    std::vector<size_t> need(T);
    for (int t = 0; t < T; ++t)
    {
        Grid& G = a.grids[a.tailStart + t];
        const Sell& S = G.sell;
        const bool coarsest = t + 1 == T;   // no sliced rows: its dense inverse is applied
        if (G.n > 65535) return;
        Host& q = h[t];
        q.nnz = coarsest ? 0 : (size_t)S.nnz;
        q.sp = coarsest ? std::vector<int>(1, 0) : download(c, S.slicePtr.p, (size_t)S.nSlices + 1);
        q.cols.assign(q.nnz, 0);
        if (!coarsest && !S.patterned)
        {
            const std::vector<int> cl = download(c, S.cols.p, q.nnz);
            for (size_t e = 0; e < cl.size(); ++e) q.cols[e] = u16(cl[e]);
        }
        else if (!coarsest)
        {
            const std::vector<uint8_t> pat = download(c, S.pattern.p, (size_t)G.n);
            const std::vector<int> po = download(c, S.patOff.p, S.patOff.n);
            for (int r = 0; r < G.n; ++r)
            {
                const int sl = r >> 6, lane = r & 63, w = (q.sp[sl + 1] - q.sp[sl]) >> 6;
                for (int k = 0; k < w; ++k)
                    q.cols[(size_t)q.sp[sl] + ((size_t)k << 6) + lane] = u16(r + po[(size_t)pat[r] * S.width + k]);
            }
        }
        if (!coarsest)
        {
            const int nc = a.grids[a.tailStart + t + 1].n;
            for (int v : download(c, G.cellStart.p, (size_t)nc + 1)) q.cs.push_back(u16(v));
            for (int v : download(c, G.cellList.p, (size_t)G.n)) q.cl.push_back(u16(v));
        }
        need[t] = al(4 * q.sp.size()) + al(2 * q.cols.size()) + al(2 * q.cs.size()) + al(2 * q.cl.size()) +
                  al(q.nnz * sm) + al((size_t)G.n * sv) + 5 * al((size_t)G.n * ss);
    }
    const size_t nL = (size_t)a.grids.back().n;
    size_t total = al(nL * sizeof(double));   // the coarsest right-hand side in FP64
    int TG = T;
    while (TG > 0 && total + need[TG - 1] <= budget)
    {
        total += need[TG - 1];
        --TG;
    }
    if (TG > T - 2) return;   // fewer than the two coarsest grids fit
    LdsTail d{};
    d.T = T;
    d.TG = TG;
    std::vector<uint8_t> blob;
    auto put = [&](const void* src, size_t bytes) {
        const size_t o = blob.size();
        blob.resize(al(o + bytes), 0);
        if (bytes) std::memcpy(blob.data() + o, src, bytes);
        return (int)o;
    };
    for (int t = TG; t < T; ++t)
    {
        LdsTailGrid& q = d.g[t];
        q.n = a.grids[a.tailStart + t].n;
        q.nnz = (int)h[t].nnz;
        q.oSlice = put(h[t].sp.data(), sizeof(int) * h[t].sp.size());
        q.oCols = put(h[t].cols.data(), 2 * h[t].cols.size());
        q.oStart = put(h[t].cs.data(), 2 * h[t].cs.size());
        q.oList = put(h[t].cl.data(), 2 * h[t].cl.size());
    }
    const size_t staticBytes = blob.size();
    size_t o = blob.size();
    auto take = [&](size_t bytes) {
        const size_t r = o;
        o = al(o + bytes);
        return (int)r;
    };
    for (int t = TG; t < T; ++t)
    {
        LdsTailGrid& q = d.g[t];
        q.oVal = take((size_t)q.nnz * sm);
        q.oDiag = take((size_t)q.n * sv);
        q.oB = take((size_t)q.n * ss);
        q.oX = take((size_t)q.n * ss);
        q.oY = take((size_t)q.n * ss);
        q.oR = take((size_t)q.n * ss);
        q.oE = take((size_t)q.n * ss);
    }
    d.oSb = take(nL * sizeof(double));
    d.oInv = -1;
    if (o + nL * nL * sizeof(double) <= budget) d.oInv = take(nL * nL * sizeof(double));
    if (o > budget) return;
    d.bytes = (int)o;
    // dynamic LDS beyond the default limit is opted into here, on the host before any launch (a capture
    // may follow); a device that refuses keeps tailK for this hierarchy
    const void* kern = f64 ? reinterpret_cast<const void*>(&tailMixK<double, double, double>)
                     : f32 ? reinterpret_cast<const void*>(&tailMixK<float, float, float>)
                           : reinterpret_cast<const void*>(&tailMixK<__half, float, __half>);
    if (hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, d.bytes) != hipSuccess)
    {
        (void)hipGetLastError();
        return;
    }
    a.tailLdsBlob.upload(blob, c.stream);
    // the copy plan: the blob, each staged grid's coefficients and diagonal, the inverse
    int words = 0;
    auto seg = [&](const void* src, int dst, size_t bytes) {
        if (!bytes) return;
        StageSeg& q = d.seg[d.nSegs++];
        q.src = src;
        q.dst = dst;
        q.first = words;
        q.words = (int)(bytes / 4);
        words += q.words;
    };
    seg(a.tailLdsBlob.p, 0, staticBytes);
    for (int t = TG; t < T; ++t)
    {
        const LdsTailGrid& q = d.g[t];
        Grid& G = a.grids[a.tailStart + t];
        if (q.nnz)
            seg(f64 ? (const void*)G.off64.p : f32 ? (const void*)G.off32.p : (const void*)G.off16.p, q.oVal,
                (size_t)q.nnz * sm);
        seg(f64 ? (const void*)G.diagV64.p : (const void*)G.diagV32.p, q.oDiag, (size_t)q.n * sv);
    }
    if (d.oInv >= 0) seg(a.coarseInv.p, d.oInv, nL * nL * sizeof(double));
    d.words = words;
    a.tailLdsDesc.upload(reinterpret_cast<const uint8_t*>(&d), sizeof(LdsTail), c.stream);
    hipCheck(hipStreamSynchronize(c.stream), "lds tail");
    a.tailLdsBytes = d.bytes;
    a.tailLdsGrid = a.tailStart + TG;
}

// grid g's pointers and sizes as the one-workgroup kernels read them, in the current precision
TailGrid tailGridOf(Amg& a, int g)
{
    Grid& G = a.grids[g];
    TailGrid t{};
    t.S = viewOf(G.sell);
    t.n = G.n;
    t.scale = a.scales.p + g;
    t.cellStart = G.cellStart.p;
    t.cellList = G.cellList.p;
    if (a.prec == Prec::F64)
    {
        t.val = G.off64.p;
        t.diag = G.diagV64.p;
        t.b = G.b64.p; t.x = G.x64.p; t.y = G.y64.p; t.r = G.r64.p; t.e = G.e64.p;
    }
    else if (a.prec == Prec::F32)
    {
        t.val = G.off32.p;
        t.diag = G.diagV32.p;
        t.b = G.b32.p; t.x = G.x32.p; t.y = G.y32.p; t.r = G.r32.p; t.e = G.e32.p;
    }
    else
    {
        t.val = G.off16.p;
        t.diag = G.diagV32.p;
        t.b = G.b16.p; t.x = G.x16.p; t.y = G.y16.p; t.r = G.r16.p; t.e = G.e16.p;
        t.xm = a.xmul.p + g;
    }
    return t;
}

void buildTail(Ctx& c, Amg& a)
{
    if (a.tailFor == (int)a.prec) return;
    a.tailFor = (int)a.prec;
    const int nG = (int)a.grids.size();
    a.tailStart = 0;
    for (int g = 1; g < nG; ++g)
        if (a.grids[g].n <= kTailCells)
        {
            a.tailStart = g;
            break;
        }
    // every grid's descriptor, for the batched whole solves (from grid 0)
    {
        std::vector<TailGrid> all;
        for (int g = 0; g < nG; ++g) all.push_back(tailGridOf(a, g));
        a.cycleDesc.upload(reinterpret_cast<const uint8_t*>(all.data()), all.size() * sizeof(TailGrid), c.stream);
    }
    if (!a.tailStart || nG - a.tailStart > kMaxTail)
    {
        a.tailStart = 0;
        return;
    }
    std::vector<TailGrid> d;
    for (int g = a.tailStart; g < nG; ++g) d.push_back(tailGridOf(a, g));
    a.tailDesc.upload(reinterpret_cast<const uint8_t*>(d.data()), d.size() * sizeof(TailGrid), c.stream);
    hipCheck(hipStreamSynchronize(c.stream), "tail descriptors");
    buildTailLds(c, a);
}

} // namespace

void amgCoarseMatrices(Ctx& c, Amg& a, const double* fineDiag, const double* fineUpper)
{
    const int nG = (int)a.grids.size();
    if (nG < 2) invalid("amg: the mesh is too small for a coarse grid");
    for (int gi = 0; gi + 1 < nG; ++gi)
    {
        Grid& g = a.grids[gi];
        Grid& gc = a.grids[gi + 1];
        const double* fd = gi == 0 ? fineDiag : g.diag64.p;
        const double* fu = gi == 0 ? fineUpper : g.upper64.p;
        galDiagK<<<blocks(gc.n), kTpb, 0, c.stream>>>(gc.n, g.cellStart.p, g.cellList.p, g.dfaceStart.p,
                                                       g.dfaceList.p, fd, fu, gc.diag64.p);
        if (gc.nFaces)
            galFaceK<<<blocks(gc.nFaces), kTpb, 0, c.stream>>>(gc.nFaces, g.faceStart.p, g.faceList.p, fu,
                                                                gc.upper64.p);
    }
    hipCheck(hipGetLastError(), "galerkin");
}

void amgGalerkin(Ctx& c, Amg& a, const double* fineDiag, const double* fineUpper, Prec prec)
{
    // the brae cycle's coarsest solve is a dense inverse in one workgroup (OpenFOAM's GAMG solves any size)
    if (a.nCoarsest > kMaxDenseCoarsest)
        throw Error(kErrUnsupported, "mfv: amg galerkin: the coarsest grid has " + std::to_string(a.nCoarsest) +
                                         " cells; the dense coarsest solve takes at most " +
                                         std::to_string(kMaxDenseCoarsest));
    a.coarseInv.resize((size_t)a.nCoarsest * a.nCoarsest);
    a.coarseWork.resize((size_t)a.nCoarsest * 2 * a.nCoarsest);
    amgCoarseMatrices(c, a, fineDiag, fineUpper);
    const int nG = (int)a.grids.size();
    // V-cycle copies of every non-coarsest grid in the working precision
    for (int gi = 0; gi + 1 < nG; ++gi)
    {
        Grid& g = a.grids[gi];
        const double* d = gi == 0 ? fineDiag : g.diag64.p;
        const double* u = gi == 0 ? fineUpper : g.upper64.p;
        const long long nnz = g.sell.nnz;
        if (prec == Prec::F64 || gi == 0)
        {
            if (g.off64.n != (size_t)nnz) zeroBuf(c, g.off64, nnz);
            if (g.nFaces)
                fillK<double><<<blocks(g.nFaces), kTpb, 0, c.stream>>>(g.nFaces, g.sell.slotLo.p, g.sell.slotHi.p, u, u,
                                                                       nullptr, g.off64.p);
            if (gi == 0 && g.diaM && g.nFaces)
                diaFillK<double><<<blocks(g.nFaces), kTpb, 0, c.stream>>>(g.nFaces, g.diaSlot.p, u, nullptr, g.dia64.p);
        }
        if (prec == Prec::F64)
        {
            g.diagV64.resize(g.n);
            hipCheck(hipMemcpyAsync(g.diagV64.p, d, sizeof(double) * g.n, hipMemcpyDeviceToDevice, c.stream), "diag");
        }
        else
        {
            g.diagV32.resize(g.n);
            castK<float><<<blocks(g.n), kTpb, 0, c.stream>>>(g.n, d, g.diagV32.p);
            if (prec == Prec::F32)
            {
                if (g.off32.n != (size_t)nnz) zeroBuf(c, g.off32, nnz);
                if (g.nFaces)
                    fillK<float><<<blocks(g.nFaces), kTpb, 0, c.stream>>>(g.nFaces, g.sell.slotLo.p, g.sell.slotHi.p, u,
                                                                          u, nullptr, g.off32.p);
                if (gi == 0 && cycleDia(g) && g.nFaces)
                {
                    if (g.dia32.n != g.dia64.n) zeroBuf(c, g.dia32, g.dia64.n);
                    diaFillK<float><<<blocks(g.nFaces), kTpb, 0, c.stream>>>(g.nFaces, g.diaSlot.p, u, nullptr,
                                                                             g.dia32.p);
                }
            }
            else
            {
                if (g.off16.n != (size_t)nnz) zeroBuf(c, g.off16, nnz);
                if (g.nFaces)
                {
                    const int gm = reduceGrid(g.nFaces);
                    double* pm = c.partials.p + 4 * (size_t)kMaxReduceBlocks;
                    maxAbsK<<<gm, kReduceBlock, 0, c.stream>>>(g.nFaces, u, pm);
                    maxPow2FinishK<<<1, kFinishBlock, 0, c.stream>>>(gm, pm, a.scales.p + gi);
                    fillK<__half><<<blocks(g.nFaces), kTpb, 0, c.stream>>>(g.nFaces, g.sell.slotLo.p, g.sell.slotHi.p,
                                                                           u, u, a.scales.p + gi, g.off16.p);
                    if (gi == 0 && cycleDia(g))
                    {
                        if (g.dia16.n != g.dia64.n) zeroBuf(c, g.dia16, g.dia64.n);
                        diaFillK<__half><<<blocks(g.nFaces), kTpb, 0, c.stream>>>(g.nFaces, g.diaSlot.p, u,
                                                                                  a.scales.p + gi, g.dia16.p);
                    }
                }
            }
        }
    }
    // coarsest: vector-precision diag (the dense solve reads FP64 face values), dense inverse
    Grid& gl = a.grids.back();
    if (prec == Prec::F64)
    {
        gl.diagV64.resize(gl.n);
        hipCheck(hipMemcpyAsync(gl.diagV64.p, gl.diag64.p, sizeof(double) * gl.n, hipMemcpyDeviceToDevice, c.stream),
                 "coarse diag");
    }
    else
    {
        gl.diagV32.resize(gl.n);
        castK<float><<<blocks(gl.n), kTpb, 0, c.stream>>>(gl.n, gl.diag64.p, gl.diagV32.p);
    }
    coarseInverseK<<<1, kTpb, 0, c.stream>>>(gl.n, gl.nFaces, gl.denseOwn.p, gl.denseNei.p, gl.diag64.p, gl.upper64.p,
                                             a.coarseWork.p, a.coarseInv.p);
    hipCheck(hipGetLastError(), "coarsest inverse");
    if (!a.allocatedVectors || a.vectorsFor != prec)
    {
        for (Grid& g : a.grids)
        {
            if (prec == Prec::F64)
                allocVectors<double>(c, g);
            else if (prec == Prec::F32)
                allocVectors<float>(c, g);
            else
                allocVectors<__half>(c, g);
        }
        a.allocatedVectors = true;
        a.vectorsFor = prec;
        a.tailFor = -1;
    }
    if (prec == Prec::F16)
    {
        // each grid's solution-vector multiplier from its FP64 diagonal
        a.xmul.resize(nG);
        a.sig.resize(2);
        double* pmx = c.partials.p + 6 * (size_t)kMaxReduceBlocks;
        double* pmn = c.partials.p + 7 * (size_t)kMaxReduceBlocks;
        for (int gi = 0; gi < nG; ++gi)
        {
            const Grid& g = a.grids[gi];
            const double* d = gi == 0 ? fineDiag : g.diag64.p;
            const int gr = reduceGrid(g.n);
            diagRangeK<<<gr, kReduceBlock, 0, c.stream>>>(g.n, d, pmx, pmn);
            xmulFinishK<<<1, kFinishBlock, 0, c.stream>>>(gr, pmx, pmn, a.xmul.p + gi);
        }
        hipCheck(hipGetLastError(), "fp16 vector multipliers");
    }
    a.prec = prec;
    buildTail(c, a);
    a.galerkinDone = true;
}

void amgStartFrom(Ctx& c, Amg& a, const double* r, const VcycleControls& vc, const int* stop)
{
    Grid& g = a.grids[0];
    const int fromZero = vc.nPre > 0 ? 1 : 0;
    if (a.prec == Prec::F64)
        startK<double, double><<<blocks(g.n), kTpb, 0, c.stream>>>(g.n, r, g.diagV64.p, g.b64.p, g.x64.p, nullptr,
                                                                    nullptr, fromZero, stop);
    else if (a.prec == Prec::F32)
        startK<float, float><<<blocks(g.n), kTpb, 0, c.stream>>>(g.n, r, g.diagV32.p, g.b32.p, g.x32.p, nullptr,
                                                                  nullptr, fromZero, stop);
    else
    {
        // the FP16 cycle's input scale: the power of two at or above max|r|, for this cycle and (until the
        // solver's update replaces it) the next
        const int gr = reduceGrid(g.n);
        double* pm = c.partials.p + 4 * (size_t)kMaxReduceBlocks;
        maxAbsK<<<gr, kReduceBlock, 0, c.stream>>>(g.n, r, pm);
        sigFinishK<<<1, kFinishBlock, 0, c.stream>>>(gr, pm, a.sig.p);
        startK<float, __half><<<blocks(g.n), kTpb, 0, c.stream>>>(g.n, r, g.diagV32.p, g.b16.p, g.x16.p, a.xmul.p,
                                                                   a.sig.p, fromZero, stop);
    }
    hipCheck(hipGetLastError(), "vcycle start");
}

void* amgVcycle(Ctx& c, Amg& a, const VcycleControls& vc, const int* stop)
{
    if (!a.galerkinDone) throw Error(kErrState, "mfv: amg_galerkin must run before a V-cycle");
    switch (a.prec)
    {
    case Prec::F64: return vcycleAt<double, double, double>(c, a, 0, vc, stop, vecs64, Scatter<double, double>{});
    case Prec::F32: return vcycleAt<float, float, float>(c, a, 0, vc, stop, vecs32, Scatter<float, float>{});
    case Prec::F16: return vcycleAt<__half, float, __half>(c, a, 0, vc, stop, vecs16, Scatter<float, __half>{});
    }
    return nullptr;
}

uint64_t amgLayoutKey(const Amg& a)
{
    uint64_t h = 0x9e3779b97f4a7c15ull;
    auto mix = [&h](uint64_t v) {
        h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
        h *= 0xff51afd7ed558ccdull;
    };
    auto ptr = [&mix](const void* p) { mix(reinterpret_cast<uintptr_t>(p)); };
    mix((uint64_t)a.prec);
    mix((uint64_t)a.tailStart);
    mix(a.grids.size());
    for (const Grid& g : a.grids)
    {
        for (const void* p : {(const void*)g.sell.slicePtr.p, (const void*)g.sell.cols.p, (const void*)g.sell.pattern.p,
                              (const void*)g.sell.patOff.p, (const void*)g.off64.p, (const void*)g.off32.p,
                              (const void*)g.off16.p, (const void*)g.diagV64.p, (const void*)g.diagV32.p,
                              (const void*)g.b64.p, (const void*)g.x64.p, (const void*)g.y64.p, (const void*)g.r64.p,
                              (const void*)g.e64.p, (const void*)g.b32.p, (const void*)g.x32.p, (const void*)g.y32.p,
                              (const void*)g.r32.p, (const void*)g.e32.p, (const void*)g.map.p,
                              (const void*)g.cellStart.p, (const void*)g.cellList.p, (const void*)g.quads.p,
                              (const void*)g.dia64.p, (const void*)g.dia32.p, (const void*)g.dia16.p})
            ptr(p);
    }
    for (const Grid& g : a.grids)
        for (const void* p : {(const void*)g.b16.p, (const void*)g.x16.p, (const void*)g.y16.p, (const void*)g.r16.p,
                              (const void*)g.e16.p})
            ptr(p);
    for (const void* p : {(const void*)a.coarseInv.p, (const void*)a.scales.p, (const void*)a.alphas.p,
                          (const void*)a.seqA.p, (const void*)a.seqB.p, (const void*)a.tailDesc.p,
                          (const void*)a.xmul.p, (const void*)a.sig.p})
        ptr(p);
    return h;
}

void amgResultToDouble(Ctx& c, Amg& a, const void* res, double* z)
{
    const int n = a.grids[0].n;
    if (a.prec == Prec::F64)
        hipCheck(hipMemcpyAsync(z, res, sizeof(double) * n, hipMemcpyDeviceToDevice, c.stream), "result copy");
    else if (a.prec == Prec::F32)
        widenK<float, float><<<blocks(n), kTpb, 0, c.stream>>>(n, static_cast<const float*>(res), nullptr, nullptr, z);
    else
        widenK<float, __half><<<blocks(n), kTpb, 0, c.stream>>>(n, static_cast<const __half*>(res), a.xmul.p,
                                                                a.sig.p, z);
    hipCheck(hipGetLastError(), "result");
}

void sellAmul64(Ctx& c, const Grid& g, const double* diag, const double* x, double* y)
{
    sellAmul64K<<<blocks(g.n), kTpb, 0, c.stream>>>(viewOf(g.sell), g.off64.p, diag, x, y);
    hipCheck(hipGetLastError(), "sell amul");
}

// The batch's solves, grouped by cycle precision, one launch of batchPcgK per group on the context's
// stream (one workgroup per solve); their work vectors on each hierarchy. Synchronises.
void amgBatchSolve(Ctx& c, std::vector<BatchSolve>& solves)
{
    for (BatchSolve& s : solves)
    {
        Amg& a = *s.a;
        if (!a.galerkinDone || a.prec != s.pc.prec)
            invalid("batch solve: the hierarchy's Galerkin values are not in the solve's precision");
        if ((int)a.grids.size() > kMaxTail || (int)a.grids.size() < 2 || a.grids.back().n > kMaxDenseCoarsest)
            throw Error(kErrUnsupported, "mfv: batch solve: the hierarchy is too deep or its coarsest grid too large");
        if (a.grids[0].n > (1 << 22)) invalid("batch solve: a system above 4 M cells belongs to the per-grid kernels");
        if (a.batchWork.n != 4 * (size_t)a.grids[0].n) a.batchWork.resize(4 * (size_t)a.grids[0].n);
    }
    const int nAll = (int)solves.size();
    if (!nAll) return;
    {
        std::vector<const Amg*> seen;
        for (const BatchSolve& s : solves) seen.push_back(s.a);
        std::sort(seen.begin(), seen.end());
        if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
            invalid("batch solve: a hierarchy appears twice in one batch (its cycle vectors are shared)");
    }
    DBuf<double> outD(2 * (size_t)nAll);
    DBuf<int> outI(2 * (size_t)nAll);
    std::vector<DBuf<uint8_t>> jobBufs;
    for (Prec prec : {Prec::F64, Prec::F32, Prec::F16})
    {
        std::vector<BatchJob> jobs;
        for (int k = 0; k < nAll; ++k)
        {
            BatchSolve& s = solves[k];
            if (s.pc.prec != prec) continue;
            Amg& a = *s.a;
            const Grid& g0 = a.grids[0];
            const size_t n = (size_t)g0.n;
            BatchJob j{};
            j.n = g0.n;
            j.T = (int)a.grids.size();
            j.grids = reinterpret_cast<const TailGrid*>(a.cycleDesc.p);
            j.S0 = viewOf(g0.sell);
            j.val0 = g0.off64.p;
            j.diag = s.diag;
            j.b = s.b;
            j.x = s.psi;
            j.r0 = a.batchWork.p;
            j.r1 = a.batchWork.p + n;
            j.p = a.batchWork.p + 2 * n;
            j.q = a.batchWork.p + 3 * n;
            j.inv = a.coarseInv.p;
            j.tol = s.pc.tolerance;
            j.relTol = s.pc.relTol;
            j.maxIter = s.pc.maxIter;
            j.minIter = s.pc.minIter;
            j.nPre = s.pc.nPre;
            j.nPost = s.pc.nPost;
            j.half = prec == Prec::F16 ? 1 : 0;
            j.flexible = s.pc.flexible ? 1 : 0;
            j.scaled = s.pc.scaleCorrection ? 1 : 0;
            j.outD = outD.p + 2 * k;
            j.outI = outI.p + 2 * k;
            jobs.push_back(j);
        }
        if (jobs.empty()) continue;
        jobBufs.emplace_back();
        jobBufs.back().upload(reinterpret_cast<const uint8_t*>(jobs.data()), jobs.size() * sizeof(BatchJob), c.stream);
        const BatchJob* dj = reinterpret_cast<const BatchJob*>(jobBufs.back().p);
        const int nj = (int)jobs.size();
        if (prec == Prec::F64)
            batchPcgK<double, double, double><<<nj, kBatchThreads, 0, c.stream>>>(dj);
        else if (prec == Prec::F32)
            batchPcgK<float, float, float><<<nj, kBatchThreads, 0, c.stream>>>(dj);
        else
            batchPcgK<__half, float, __half><<<nj, kBatchThreads, 0, c.stream>>>(dj);
        hipCheck(hipGetLastError(), "batch solve");
    }
    const std::vector<double> hd = outD.download(c.stream);
    const std::vector<int> hi = outI.download(c.stream);
    for (int k = 0; k < nAll; ++k)
    {
        PcgPerf& p = solves[k].perf;
        p.initialResidual = hd[2 * k];
        p.finalResidual = hd[2 * k + 1];
        p.nIterations = hi[2 * k];
        p.converged = hi[2 * k + 1] != 0;
    }
}

} // namespace mfv
