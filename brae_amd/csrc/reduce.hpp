// Deterministic wave64 reductions and the XCD-aware row partition for gfx950.
//
// Partition. A kernel over n rows runs reduceGrid(n) workgroups of kReduceBlock threads; workgroup w
// owns the contiguous chunk of rows [w*chunk, (w+1)*chunk) (chunk a multiple of kReduceBlock, a
// function of n only). The hardware hands workgroups to the eight XCDs round-robin, so xcdBlock()
// renumbers them such that XCD k works through one contiguous eighth of the rows in order: a row's
// neighbours (c +- 1, c +- nx, ...) were just touched by the same XCD and sit in its L2.
//
// Reductions. Each workgroup sums its rows in ascending order per thread, combines its four
// wavefronts by a fixed xor butterfly and a fixed wave order, and stores one partial; a second
// one-workgroup kernel sums the partials in a fixed order and applies the caller's post step. The
// result depends on n and the data only -- the reference's two-stage reductions (reductions.cu:1-21)
// -- and no kernel needs a device-scope fence (which on MI355X writes back the L2).
#pragma once

#include "mfv_internal.hpp"

namespace mfv {

inline int reduceGrid(int n)
{
    int b = (n + kReduceBlock * 2 - 1) / (kReduceBlock * 2);
    if (b < 1) b = 1;
    if (b > kMaxReduceBlocks) b = kMaxReduceBlocks;
    return b;
}

// a partition of `rows` rows per thread
inline int reduceGridRows(int n, int rows)
{
    const long long per = (long long)kReduceBlock * rows;
    long long b = ((long long)n + per - 1) / per;
    if (b < 1) b = 1;
    if (b > kMaxReduceBlocks) b = kMaxReduceBlocks;
    return (int)b;
}

// rows per workgroup for a reduceGrid partition (a multiple of kReduceBlock)
__host__ __device__ __forceinline__ int chunkRows(int n, int grid)
{
    const int per = (n + grid - 1) / grid;
    return (per + kReduceBlock - 1) / kReduceBlock * kReduceBlock;
}

// workgroup index renumbered so each XCD processes a contiguous run of workgroups
__device__ __forceinline__ int xcdBlock()
{
    const int G = gridDim.x, b = blockIdx.x;
    const int q = G >> 3, r = G & 7, x = b & 7, i = b >> 3;
    return x * q + (x < r ? x : r) + i;
}

__device__ __forceinline__ double waveSum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Max over a workgroup of N threads (exact, so any order); the result is valid in thread 0. It starts
// from the first wave's maximum, so negative values (a min taken as -max(-v)) come out right, and ends
// with a barrier, as blockSum does: the LDS slots are free again for the caller's next reduction only
// once thread 0 has read them.
template <int N = kReduceBlock>
__device__ __forceinline__ double blockMax(double v, double* lds)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    double m = 0.0;
    if (threadIdx.x == 0)
    {
        m = lds[0];
        for (int w = 1; w < N / 64; ++w) m = fmax(m, lds[w]);
    }
    __syncthreads();
    return m;
}

// Sum over a workgroup of N threads; the result is valid in thread 0.
template <int N = kReduceBlock>
__device__ __forceinline__ double blockSum(double v, double* lds)
{
    v = waveSum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
    {
#pragma unroll
        for (int w = 0; w < N / 64; ++w) s += lds[w];
    }
    __syncthreads();
    return s;
}

// Store this workgroup's partial (thread 0 holds the block sum).
__device__ __forceinline__ void storePartial(double blockTotal, double* partials)
{
    if (threadIdx.x == 0) partials[blockIdx.x] = blockTotal;
}

// One finishing thread's share of nParts partials (i = tid, tid + kFinishBlock, ...): four running sums in
// a fixed rotation, combined in a fixed order. With Batch 32 a large partition (the 20 M-row kernels leave
// tens of thousands of partials) is read 32 and then 16 loads at a time, all issued before the first add,
// so the finish waits for a few memory round trips instead of one per four partials; the adds keep the
// rotation, so every Batch gives the same bits. Batch 4 holds no array of loads (for kernels that inline
// a small finish and must keep their registers).
template <int Batch>
__device__ __forceinline__ double finishShareB(int nParts, const double* __restrict__ pa, int tid)
{
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int i = tid;
    if constexpr (Batch >= 32)
        for (; i + 31 * kFinishBlock < nParts; i += 32 * kFinishBlock)
        {
            double v[32];
#pragma unroll
            for (int u = 0; u < 32; ++u) v[u] = pa[i + u * kFinishBlock];
#pragma unroll
            for (int u = 0; u < 32; u += 4)
            {
                a0 += v[u];
                a1 += v[u + 1];
                a2 += v[u + 2];
                a3 += v[u + 3];
            }
        }
    if constexpr (Batch >= 16)
        for (; i + 15 * kFinishBlock < nParts; i += 16 * kFinishBlock)
        {
            double v[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) v[u] = pa[i + u * kFinishBlock];
#pragma unroll
            for (int u = 0; u < 16; u += 4)
            {
                a0 += v[u];
                a1 += v[u + 1];
                a2 += v[u + 2];
                a3 += v[u + 3];
            }
        }
    for (; i + 3 * kFinishBlock < nParts; i += 4 * kFinishBlock)
    {
        a0 += pa[i];
        a1 += pa[i + kFinishBlock];
        a2 += pa[i + 2 * kFinishBlock];
        a3 += pa[i + 3 * kFinishBlock];
    }
    for (; i < nParts; i += kFinishBlock) a0 += pa[i];
    return (a0 + a1) + (a2 + a3);
}
__device__ __forceinline__ double finishShare(int nParts, const double* __restrict__ pa, int tid)
{
    return finishShareB<4>(nParts, pa, tid);
}
// the share of a kFinishBlock-thread finishing kernel
__device__ __forceinline__ double finishShare(int nParts, const double* __restrict__ pa)
{
    return finishShareB<32>(nParts, pa, (int)threadIdx.x);
}

// finishShare(nParts, pa) and finishMax(nParts, pm, 0) together, the two arrays' loads issued in the same
// rounds (sixteen of each in flight); the sum keeps finishShare's rotation, so it is the same bits.
__device__ __forceinline__ void finishShareMax(int nParts, const double* __restrict__ pa,
                                               const double* __restrict__ pm, double& share, double& mx)
{
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, m = 0.0;
    int i = threadIdx.x;
    for (; i + 15 * kFinishBlock < nParts; i += 16 * kFinishBlock)
    {
        double v[16], w[16];
#pragma unroll
        for (int u = 0; u < 16; ++u)
        {
            v[u] = pa[i + u * kFinishBlock];
            w[u] = pm[i + u * kFinishBlock];
        }
#pragma unroll
        for (int u = 0; u < 16; u += 4)
        {
            a0 += v[u];
            a1 += v[u + 1];
            a2 += v[u + 2];
            a3 += v[u + 3];
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) m = fmax(m, w[u]);
    }
    for (; i + 3 * kFinishBlock < nParts; i += 4 * kFinishBlock)
    {
        a0 += pa[i];
        a1 += pa[i + kFinishBlock];
        a2 += pa[i + 2 * kFinishBlock];
        a3 += pa[i + 3 * kFinishBlock];
        m = fmax(fmax(fmax(fmax(m, pm[i]), pm[i + kFinishBlock]), pm[i + 2 * kFinishBlock]), pm[i + 3 * kFinishBlock]);
    }
    for (; i < nParts; i += kFinishBlock)
    {
        a0 += pa[i];
        m = fmax(m, pm[i]);
    }
    share = (a0 + a1) + (a2 + a3);
    mx = m;
}

// One finishing thread's maximum (Min: minimum) of nParts partials from m (i = threadIdx.x, + kFinishBlock,
// ...), eight loads in flight per round (a maximum is exact in any order).
template <bool Min = false>
__device__ __forceinline__ double finishMax(int nParts, const double* __restrict__ p, double m)
{
    int i = threadIdx.x;
    for (; i + 7 * kFinishBlock < nParts; i += 8 * kFinishBlock)
    {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = p[i + u * kFinishBlock];
#pragma unroll
        for (int u = 0; u < 8; ++u) m = Min ? fmin(m, v[u]) : fmax(m, v[u]);
    }
    for (; i < nParts; i += kFinishBlock) m = Min ? fmin(m, p[i]) : fmax(m, p[i]);
    return m;
}

// finish2K's two sums computed by one kReduceBlock-thread workgroup, with finish2K's exact arithmetic: each
// thread plays four of the kFinishBlock finishing threads (t, t + kReduceBlock, ...), whose shares it reduces
// over its wave by the same butterfly; thread 0 adds the sixteen wave sums in finish2K's order. Lets a
// kernel that needs a small reduction's result (a few hundred partials) finish it itself, in every
// workgroup, instead of a one-workgroup kernel between two launches. Valid in thread 0; all threads call.
__device__ __forceinline__ void finish2InBlock(int nParts, const double* __restrict__ pa, const double* __restrict__ pb,
                                               double* lds, double& sa, double& sb)
{
    static_assert(kFinishBlock == 4 * kReduceBlock, "four finishing threads per workgroup thread");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
        const int v = (int)threadIdx.x + k * kReduceBlock;
        const double a = waveSum(finishShare(nParts, pa, v));
        const double b = waveSum(finishShare(nParts, pb, v));
        if (lane == 0)
        {
            lds[wave + 4 * k] = a;
            lds[16 + wave + 4 * k] = b;
        }
    }
    __syncthreads();
    sa = 0.0;
    sb = 0.0;
    if (threadIdx.x == 0)
    {
#pragma unroll
        for (int w = 0; w < kFinishBlock / 64; ++w) sa += lds[w];
#pragma unroll
        for (int w = 0; w < kFinishBlock / 64; ++w) sb += lds[16 + w];
    }
    __syncthreads();
}

// Stage two: one workgroup sums nParts partials in a fixed order and hands the total to `post`.
template <typename Post>
__global__ __launch_bounds__(kFinishBlock) void finishK(int nParts, const double* __restrict__ pa, Post post,
                                                        const int* __restrict__ stop)
{
    if (stop && *stop) return;
    __shared__ double lds[kFinishBlock / 64];
    const double s = blockSum<kFinishBlock>(finishShare(nParts, pa), lds);
    if (threadIdx.x == 0) post(s);
}

template <typename Post>
__global__ __launch_bounds__(kFinishBlock) void finish2K(int nParts, const double* __restrict__ pa,
                                                         const double* __restrict__ pb, Post post,
                                                         const int* __restrict__ stop)
{
    if (stop && *stop) return;
    __shared__ double lds[kFinishBlock / 64];
    // both shares' loads issue before the first barrier
    const double ha = finishShare(nParts, pa), hb = finishShare(nParts, pb);
    const double sa = blockSum<kFinishBlock>(ha, lds);
    const double sb = blockSum<kFinishBlock>(hb, lds);
    if (threadIdx.x == 0) post(sa, sb);
}

// Sequential (OpenFOAM-order) completion of a reduction whose per-element terms were stored: one
// thread sums them in ascending index, the host loop's order, and applies the same post step.
template <typename Post>
__global__ void seqFinishK(int n, const double* __restrict__ a, Post post, const int* __restrict__ stop)
{
    if (threadIdx.x || blockIdx.x) return;
    if (stop && *stop) return;
    double sa = 0.0;
    for (int i = 0; i < n; ++i) sa += a[i];
    post(sa);
}

template <typename Post>
__global__ void seqFinish2K(int n, const double* __restrict__ a, const double* __restrict__ b, Post post,
                            const int* __restrict__ stop)
{
    if (threadIdx.x || blockIdx.x) return;
    if (stop && *stop) return;
    double sa = 0.0, sb = 0.0;
    for (int i = 0; i < n; ++i)
    {
        sa += a[i];
        sb += b[i];
    }
    post(sa, sb);
}

struct StoreTo
{
    double* out;
    __device__ void operator()(double s) const { *out = s; }
};

// Plane-slab partition of a grid whose rows form planes of `plane` rows (a fixed stencil's largest
// offset; the last plane may be partial). XCD k -- workgroup b runs on XCD b % 8 -- takes the k-th
// eighth (slab) of every plane, plane after plane, and each of its workgroups the next rpt*kReduceBlock
// rows of that sequence, thread t rows t, t + kReduceBlock, ... A row's neighbour one plane away was
// then touched by the same XCD one slab earlier, still in its L2, where the contiguous partition has it
// an eighth of the whole grid earlier. Every row falls to exactly one thread. plane >= 8 * 64.
struct SlabPart
{
    int n, plane, rpt;
};

constexpr int kSlabMinPlane = 8 * 64;

// workgroups of a slab partition with rpt rows per thread
inline int slabBlocks(int n, int plane, int rpt)
{
    const long long nPl = ((long long)n + plane - 1) / plane;
    const long long slabMax = (plane + 7) / 8;
    const long long per = (long long)kReduceBlock * rpt;
    return (int)(8 * ((nPl * slabMax + per - 1) / per));
}

// the fewest rows per thread that keep the workgroups within kMaxReduceBlocks partials
inline SlabPart slabPart(int n, int plane)
{
    int rpt = 1;
    while (slabBlocks(n, plane, rpt) > kMaxReduceBlocks) ++rpt;
    return SlabPart{n, plane, rpt};
}

// this thread's r-th row (r < rpt), or -1
__device__ __forceinline__ int slabRow(const SlabPart& m, int r)
{
    const int k = blockIdx.x & 7;
    const int s0 = (int)((long long)k * m.plane >> 3), s1 = (int)((long long)(k + 1) * m.plane >> 3);
    const int slab = s1 - s0;
    const int p = ((int)(blockIdx.x >> 3) * m.rpt + r) * kReduceBlock + (int)threadIdx.x;
    const int z = p / slab;
    const long long row = (long long)z * m.plane + s0 + (p - z * slab);
    return row < m.n ? (int)row : -1;
}

// The cell of a per-cell gather kernel's thread, or -1. Slab: the plane-slab partition of the matrix's
// band (a cell's neighbour one band back -- lower[], psi[] of its losort run -- was read by the same XCD
// one slab earlier and is still in that XCD's L2); else the XCD-contiguous order.
// With Split > 1 each XCD's slab is cut into that many column parts, and the XCD walks each part through
// every plane before the next: the reuse distance of a band-back neighbour shrinks to one part, at the
// price of the cells next to a part boundary, whose neighbours across it were read a whole walk earlier.
// Measured at 20 M cells: two parts speed up linearUpwindCorrK (1762 -> 1662 us) and slow the other
// gathers by 1-5 %, four parts are slower for all of them.
template <bool Slab, int Split = 1>
__device__ __forceinline__ int gatherCell(int n, const SlabPart& part)
{
    if (!Slab)
    {
        const int c = xcdBlock() * kReduceBlock + threadIdx.x;
        return c < n ? c : -1;
    }
    const int k = blockIdx.x & 7;
    const int s0 = (int)((long long)k * part.plane >> 3), s1 = (int)((long long)(k + 1) * part.plane >> 3);
    const int slab = s1 - s0, sub = (slab + Split - 1) / Split;
    const int nPl = (n + part.plane - 1) / part.plane;
    const long long p = (long long)(blockIdx.x >> 3) * kReduceBlock + threadIdx.x;
    const long long per = (long long)nPl * sub;
    const int col = (int)(p / per);
    if (col >= Split) return -1;
    const long long q = p - col * per;
    const int z = (int)(q / sub), x = col * sub + (int)(q - (long long)z * sub);
    if (x >= slab) return -1;
    const long long row = (long long)z * part.plane + s0 + x;
    return row < n ? (int)row : -1;
}

// workgroups of the gather partition
inline int gatherBlocks(int n, int plane, int split = 1)
{
    const long long nPl = ((long long)n + plane - 1) / plane;
    const long long slabMax = (plane + 7) / 8, sub = (slabMax + split - 1) / split;
    return (int)(8 * ((split * nPl * sub + kReduceBlock - 1) / kReduceBlock));
}

// the partition of a mesh's cells for gather kernels: slabs of the band when the cells span several
// bands of at least kSlabMinPlane cells, one row per thread
inline SlabPart gatherPart(const Mesh& m)
{
    if (m.band >= kSlabMinPlane && m.band < m.nCells) return SlabPart{m.nCells, m.band, 1};
    return SlabPart{m.nCells, 0, 1};
}

// Row loop over this workgroup's chunk of a reduceGrid partition, XCD-ordered.
#define MFV_ROWS(n, c)                                                                                            \
    const int mfvChunk_ = chunkRows((n), gridDim.x);                                                             \
    const int mfvBegin_ = xcdBlock() * mfvChunk_;                                                                 \
    const int mfvEnd_ = (mfvBegin_ + mfvChunk_ < (n)) ? mfvBegin_ + mfvChunk_ : (n);                              \
    for (int c = mfvBegin_ + (int)threadIdx.x; c < mfvEnd_; c += kReduceBlock)

} // namespace mfv
