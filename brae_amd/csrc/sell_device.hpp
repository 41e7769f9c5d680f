// Device-side helpers shared by the kernels that read a grid's sliced rows (amg.hip, pcg.hip).
#pragma once

#include "amg.hpp"

namespace mfv {

template <typename T>
__device__ __forceinline__ T safeDiag(T d)
{
    const T fl = T(sizeof(T) == 4 ? 1e-30 : 1e-300);
    if (d > fl) return d;
    if (d < -fl) return d;
    return d < T(0) ? -fl : fl;
}

// A load of data a pass reads once (coefficients, explicit columns, row patterns, diagonals):
// nontemporal, so the stream does not push the gathered vectors -- reread by the rows a plane of
// cells away -- out of the XCD's L2.
template <typename T>
__device__ __forceinline__ T ldStream(const T* p)
{
    return __builtin_nontemporal_load(p);
}
template <>
__device__ __forceinline__ __half ldStream<__half>(const __half* p)
{
    return __ushort_as_half(__builtin_nontemporal_load(reinterpret_cast<const unsigned short*>(p)));
}
template <>
__device__ __forceinline__ float4 ldStream<float4>(const float4* p)
{
    const float4 v = *p;
    return v;
}

// a stored coefficient in the arithmetic precision Tv
template <typename Tv, typename Tm>
__device__ __forceinline__ Tv cvt(Tm v) { return (Tv)v; }
template <>
__device__ __forceinline__ float cvt<float, __half>(__half v) { return __half2float(v); }

// Sliced rows of one grid, as kernel arguments. With diaM (2 or 3) the values a kernel gets are the
// grid's diagonal copy instead (value of face (r, r + diaOff[k]) at k * diaPad + r), read by diaOffSum.
struct SellView
{
    int n;
    const int* slicePtr;
    const int* cols;
    const uint8_t* pattern;
    const int* patOff;
    int width;
    int diaM;
    int diaPad;
    int diaOff[kDiaMax];
};

inline SellView viewOf(const Sell& s)
{
    return SellView{s.n, s.slicePtr.p, s.patterned ? nullptr : s.cols.p, s.patterned ? s.pattern.p : nullptr,
                    s.patterned ? s.patOff.p : nullptr, s.width, 0, 0, {}};
}

// sum_k A'_ck * X(col_k) over row c's entries in face order, in the vector precision. X is a functor
// returning the operand at a column (plain x, x + prolonged correction, a direction vector built on the
// fly). Rows are handed out with each wavefront's 64 lanes on one slice (row c in lane c & 63), so the
// slice's start and width are wave-uniform: they are read once and kept in scalar registers, and the
// width selects straight-line code that issues every coefficient (and column) load of a group of up to
// eight entries before the gathers, then sums in entry order. Padded slots hold a zero coefficient and
// the row's own column.
template <int W, bool Pat, typename Tm, typename Tv, typename XF>
__device__ __forceinline__ Tv rowGroup(const SellView& S, int c, int lane, int base, const int* off,
                                       const Tm* __restrict__ val, XF X, Tv acc)
{
    Tv a[W];
    int col[W];
#pragma unroll
    for (int k = 0; k < W; ++k)
    {
        const int e = base + (k << 6) + lane;
        a[k] = cvt<Tv>(ldStream(val + e));
        col[k] = Pat ? c + off[k] : ldStream(S.cols + e);
    }
    Tv xv[W];
#pragma unroll
    for (int k = 0; k < W; ++k) xv[k] = X(col[k]);
#pragma unroll
    for (int k = 0; k < W; ++k) acc += a[k] * xv[k];
    return acc;
}

template <bool Pat, typename Tm, typename Tv, typename XF>
__device__ __forceinline__ Tv rowEntries(const SellView& S, int c, int lane, int p0, int w, const int* off,
                                         const Tm* __restrict__ val, XF X, Tv acc)
{
    int k0 = 0;
    for (; k0 + 8 <= w; k0 += 8)
        acc = rowGroup<8, Pat, Tm, Tv>(S, c, lane, p0 + (k0 << 6), Pat ? off + k0 : off, val, X, acc);
    const int* o = Pat ? off + k0 : off;
    const int base = p0 + (k0 << 6);
    switch (w - k0)
    {
    case 1: return rowGroup<1, Pat, Tm, Tv>(S, c, lane, base, o, val, X, acc);
    case 2: return rowGroup<2, Pat, Tm, Tv>(S, c, lane, base, o, val, X, acc);
    case 3: return rowGroup<3, Pat, Tm, Tv>(S, c, lane, base, o, val, X, acc);
    case 4: return rowGroup<4, Pat, Tm, Tv>(S, c, lane, base, o, val, X, acc);
    case 5: return rowGroup<5, Pat, Tm, Tv>(S, c, lane, base, o, val, X, acc);
    case 6: return rowGroup<6, Pat, Tm, Tv>(S, c, lane, base, o, val, X, acc);
    case 7: return rowGroup<7, Pat, Tm, Tv>(S, c, lane, base, o, val, X, acc);
    default: return acc;
    }
}

// The same sum on a grid's diagonal copy (M offsets): row c's lower entries at columns c - off[M-1] ..
// c - off[0], then its upper ones at c + off[0] .. c + off[M-1] -- its face order -- with a zero
// coefficient where a face is absent (a column outside the grid reads the row itself). No row pattern or
// column is loaded: every load's address follows from c, so all of them issue at once. The upper
// coefficients stream (nontemporal); a lower one is the upper coefficient of the row above it, reread
// through the cache.
template <int M, typename Tm, typename Tv, typename XF>
__device__ __forceinline__ Tv diaOffSum(const SellView& S, int c, const Tm* __restrict__ val, XF X, Tv acc)
{
    Tv a[2 * M], xv[2 * M];
    int col[2 * M];
#pragma unroll
    for (int k = 0; k < M; ++k)
    {
        const int kk = M - 1 - k;
        const int j = c - S.diaOff[kk];
        const bool in = j >= 0;
        col[k] = in ? j : c;
        const Tv v = cvt<Tv>(val[(size_t)kk * S.diaPad + col[k]]);
        a[k] = in ? v : Tv(0);
    }
#pragma unroll
    for (int k = 0; k < M; ++k)
    {
        const int j = c + S.diaOff[k];
        col[M + k] = j < S.n ? j : c;
        a[M + k] = cvt<Tv>(ldStream(val + (size_t)k * S.diaPad + c));
    }
#pragma unroll
    for (int k = 0; k < 2 * M; ++k) xv[k] = X(col[k]);
#pragma unroll
    for (int k = 0; k < 2 * M; ++k) acc += a[k] * xv[k];
    return acc;
}

template <typename Tm, typename Tv, typename XF>
__device__ __forceinline__ Tv rowOffSum(const SellView& S, int c, const Tm* __restrict__ val, XF X, Tv acc)
{
    if (S.diaM == 3) return diaOffSum<3, Tm, Tv>(S, c, val, X, acc);
    if (S.diaM == 2) return diaOffSum<2, Tm, Tv>(S, c, val, X, acc);
    const int s = __builtin_amdgcn_readfirstlane(c >> 6), lane = c & 63;
    const int p0 = __builtin_amdgcn_readfirstlane(S.slicePtr[s]);
    const int w = __builtin_amdgcn_readfirstlane((S.slicePtr[s + 1] - p0) >> 6);
    if (S.pattern)
        return rowEntries<true, Tm, Tv>(S, c, lane, p0, w, S.patOff + (int)ldStream(S.pattern + c) * S.width, val, X,
                                        acc);
    return rowEntries<false, Tm, Tv>(S, c, lane, p0, w, nullptr, val, X, acc);
}

// A x at row c: D x + sum (FP64/FP32: accumulated term by term after D x, the reference's order);
// FP16: D x + s * sum(A' x) with A' = A/s stored.
// (dOut: the diagonal it loaded, for a caller that also divides by it)
template <typename Tm, typename Tv, typename XF>
__device__ __forceinline__ Tv rowAx(const SellView& S, int c, const Tm* __restrict__ val, const Tv* __restrict__ diag,
                                    const double* __restrict__ scale, XF X, Tv* dOut = nullptr)
{
    const Tv xc = X(c);
    const Tv dc = ldStream(diag + c);
    if (dOut) *dOut = dc;
    if constexpr (sizeof(Tm) == 2)
    {
        const Tv off = rowOffSum<Tm, Tv>(S, c, val, X, Tv(0));
        return dc * xc + (Tv)(*scale) * off;
    }
    else
    {
        return rowOffSum<Tm, Tv>(S, c, val, X, dc * xc);
    }
}

template <typename Tv>
struct PlainX
{
    const Tv* __restrict__ x;
    __device__ Tv operator()(int j) const { return x[j]; }
};

// A V-cycle vector of one grid, stored as Ts with arithmetic in Tv. The FP64 and FP32 cycles store
// their arithmetic type as is. The FP16 cycle stores FP16: its input is normalised by a power of two
// (the cycle is positively homogeneous, so that is exact), right-hand-side vectors (b, r, A e) are then
// O(1) on every grid, and solution vectors (x, e), which scale like b/D, are stored multiplied by the
// grid's power-of-two typical diagonal `mul`; stores saturate at the FP16 range. Usable as the row
// engine's operand functor.
template <typename Tv, typename Ts>
struct CycVec
{
    Ts* __restrict__ p;
    __device__ CycVec(const Ts* q, const float*) : p(const_cast<Ts*>(q)) {}
    __device__ CycVec(const Ts* q, float, float) : p(const_cast<Ts*>(q)) {}
    __device__ Tv operator()(int i) const { return p[i]; }
    __device__ void set(int i, Tv v) const { p[i] = v; }
};

template <>
struct CycVec<float, __half>
{
    __half* __restrict__ p;
    float m, im;
    __device__ CycVec(const __half* q, const float* mul)
        : p(const_cast<__half*>(q)), m(mul ? *mul : 1.0f), im(mul ? 1.0f / *mul : 1.0f)
    {
    }
    // the multiplier and its reciprocal already read (mul ? *mul : 1, mul ? 1 / *mul : 1)
    __device__ CycVec(const __half* q, float mul, float inv) : p(const_cast<__half*>(q)), m(mul), im(inv) {}
    __device__ float operator()(int i) const { return __half2float(p[i]) * im; }
    __device__ void set(int i, float v) const { p[i] = __float2half(fminf(fmaxf(v * m, -65504.0f), 65504.0f)); }
};


} // namespace mfv
