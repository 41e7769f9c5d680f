// Deterministic wave64 reductions for gfx950.
//
// Every reduction in the library uses ONE partition of [0, n) that depends on n only: reduceGrid(n)
// workgroups of kReduceBlock threads, each thread summing its grid-stride elements in ascending
// order, each workgroup combining its four wavefronts by a fixed xor-butterfly and a fixed
// wave order, and the last workgroup to finish (a ticket counter) summing the per-workgroup
// partials in ascending workgroup order. The bits of a result therefore depend on n and the
// data, never on scheduling -- the same guarantee the reference's two-stage reductions give
// (reductions.cu:1-21), with the second stage folded into the first launch.
#pragma once

#include "mfv_internal.hpp"

namespace mfv {

inline int reduceGrid(int n)
{
    int b = (n + kReduceBlock * 4 - 1) / (kReduceBlock * 4);
    if (b < 1) b = 1;
    if (b > kMaxReduceBlocks) b = kMaxReduceBlocks;
    return b;
}

__device__ __forceinline__ double waveSum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum over the workgroup (kReduceBlock threads); the result is valid in thread 0.
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
        for (int w = 0; w < kReduceBlock / 64; ++w) s += lds[w];
    }
    __syncthreads();
    return s;
}

// Stage two, run by every workgroup after its partial is stored: the last one to arrive sums the
// partials in ascending order and writes *out (through `post` so callers can scale or combine).
// Returns true in the workgroup that finished the reduction.
template <typename Post>
__device__ __forceinline__ bool finishReduction(double partial, double* partials, unsigned int* ticket, double* lds,
                                                Post post)
{
    __shared__ bool last;
    if (threadIdx.x == 0)
    {
        partials[blockIdx.x] = partial;
        __threadfence();
        const unsigned int t = atomicAdd(ticket, 1u);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    double v = 0.0;
    for (int i = threadIdx.x; i < (int)gridDim.x; i += kReduceBlock) v += partials[i];
    const double s = blockSum(v, lds);
    if (threadIdx.x == 0)
    {
        post(s);
        *ticket = 0u;
    }
    return true;
}

// Two sums reduced together (one ticket): the last workgroup hands both totals to `post`.
template <typename Post>
__device__ __forceinline__ bool finishReduction2(double pa, double pb, double* partialsA, double* partialsB,
                                                 unsigned int* ticket, double* lds, Post post)
{
    __shared__ bool last;
    if (threadIdx.x == 0)
    {
        partialsA[blockIdx.x] = pa;
        partialsB[blockIdx.x] = pb;
        __threadfence();
        const unsigned int t = atomicAdd(ticket, 1u);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    double va = 0.0, vb = 0.0;
    for (int i = threadIdx.x; i < (int)gridDim.x; i += kReduceBlock)
    {
        va += partialsA[i];
        vb += partialsB[i];
    }
    const double sa = blockSum(va, lds);
    const double sb = blockSum(vb, lds);
    if (threadIdx.x == 0)
    {
        post(sa, sb);
        *ticket = 0u;
    }
    return true;
}

struct StoreTo
{
    double* out;
    __device__ void operator()(double s) const { *out = s; }
};

} // namespace mfv
