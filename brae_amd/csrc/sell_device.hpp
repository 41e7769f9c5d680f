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

// a stored coefficient in the arithmetic precision Tv
template <typename Tv, typename Tm>
__device__ __forceinline__ Tv cvt(Tm v) { return (Tv)v; }
template <>
__device__ __forceinline__ float cvt<float, __half>(__half v) { return __half2float(v); }

// Sliced rows of one grid, as kernel arguments.
struct SellView
{
    int n;
    const int* slicePtr;
    const int* cols;
    const uint8_t* pattern;
    const int* patOff;
    int width;
};

inline SellView viewOf(const Sell& s)
{
    return SellView{s.n, s.slicePtr.p, s.patterned ? nullptr : s.cols.p, s.patterned ? s.pattern.p : nullptr,
                    s.patterned ? s.patOff.p : nullptr, s.width};
}

// sum_k A'_ck * X(col_k) over row c's entries in face order, in the vector precision. X is a functor
// returning the operand at a column (plain x, or x + prolonged coarse correction). Rows of up to
// kRowRegs entries load every coefficient and column first and then gather, so a row's loads are all in
// flight together; the sum is still taken in entry order.
constexpr int kRowRegs = 8;

template <typename Tm, typename Tv, typename XF>
__device__ __forceinline__ Tv rowOffSum(const SellView& S, int c, const Tm* __restrict__ val, XF X, Tv acc)
{
    const int s = c >> 6, lane = c & 63;
    const int p0 = S.slicePtr[s];
    const int w = (S.slicePtr[s + 1] - p0) >> 6;
    if (w <= kRowRegs)
    {
        Tv a[kRowRegs];
        int col[kRowRegs];
        if (S.pattern)
        {
            const int* off = S.patOff + (int)S.pattern[c] * S.width;
#pragma unroll
            for (int k = 0; k < kRowRegs; ++k)
                if (k < w)
                {
                    a[k] = cvt<Tv>(val[p0 + (k << 6) + lane]);
                    col[k] = c + off[k];
                }
        }
        else
        {
#pragma unroll
            for (int k = 0; k < kRowRegs; ++k)
                if (k < w)
                {
                    const int e = p0 + (k << 6) + lane;
                    a[k] = cvt<Tv>(val[e]);
                    col[k] = S.cols[e];
                }
        }
        Tv xv[kRowRegs];
#pragma unroll
        for (int k = 0; k < kRowRegs; ++k)
            if (k < w) xv[k] = X(col[k]);
#pragma unroll
        for (int k = 0; k < kRowRegs; ++k)
            if (k < w) acc += a[k] * xv[k];
        return acc;
    }
    if (S.pattern)
    {
        const int* off = S.patOff + (int)S.pattern[c] * S.width;
        for (int k = 0; k < w; ++k)
        {
            const Tv a = cvt<Tv>(val[p0 + (k << 6) + lane]);
            acc += a * X(c + off[k]);
        }
    }
    else
    {
        for (int k = 0; k < w; ++k)
        {
            const int e = p0 + (k << 6) + lane;
            const Tv a = cvt<Tv>(val[e]);
            acc += a * X(S.cols[e]);
        }
    }
    return acc;
}

// A x at row c: D x + sum (FP64/FP32: accumulated term by term after D x, the reference's order);
// FP16: D x + s * sum(A' x) with A' = A/s stored.
template <typename Tm, typename Tv, typename XF>
__device__ __forceinline__ Tv rowAx(const SellView& S, int c, const Tm* __restrict__ val, const Tv* __restrict__ diag,
                                    const double* __restrict__ scale, XF X)
{
    const Tv xc = X(c);
    if constexpr (sizeof(Tm) == 2)
    {
        const Tv off = rowOffSum<Tm, Tv>(S, c, val, X, Tv(0));
        return diag[c] * xc + (Tv)(*scale) * off;
    }
    else
    {
        return rowOffSum<Tm, Tv>(S, c, val, X, diag[c] * xc);
    }
}

template <typename Tv>
struct PlainX
{
    const Tv* __restrict__ x;
    __device__ Tv operator()(int j) const { return x[j]; }
};

template <typename Tv>
struct ProlongedX
{
    const Tv* __restrict__ x;
    const int* __restrict__ map;
    const Tv* __restrict__ xc;
    __device__ Tv operator()(int j) const { return x[j] + xc[map[j]]; }
};


} // namespace mfv
