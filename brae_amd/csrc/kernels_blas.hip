// FP64 reductions and BLAS-1 on the device: the deterministic dot and sum|x| the PCG recurrence
// and OpenFOAM's residual norm use (reference: reductions.cu:112-147,222-250 dotKernel/sumMagKernel/
// finalSumKernel; the sequential OpenFOAM-order mode is its BRAE_DEVICE_OF_REDUCE, :36-56).
#include "reduce.hpp"

namespace mfv {

namespace {

__global__ __launch_bounds__(kReduceBlock) void dotK(int n, const double* __restrict__ a, const double* __restrict__ b,
                                                     double* __restrict__ partials, unsigned int* __restrict__ ticket,
                                                     double* __restrict__ out)
{
    __shared__ double lds[kReduceBlock / 64];
    double v = 0.0;
    for (int i = blockIdx.x * kReduceBlock + threadIdx.x; i < n; i += gridDim.x * kReduceBlock) v += a[i] * b[i];
    const double s = blockSum(v, lds);
    finishReduction(s, partials, ticket, lds, StoreTo{out});
}

__global__ __launch_bounds__(kReduceBlock) void sumMagK(int n, const double* __restrict__ a,
                                                        double* __restrict__ partials,
                                                        unsigned int* __restrict__ ticket, double* __restrict__ out)
{
    __shared__ double lds[kReduceBlock / 64];
    double v = 0.0;
    for (int i = blockIdx.x * kReduceBlock + threadIdx.x; i < n; i += gridDim.x * kReduceBlock) v += fabs(a[i]);
    const double s = blockSum(v, lds);
    finishReduction(s, partials, ticket, lds, StoreTo{out});
}

// OpenFOAM's own order: one thread, ascending index (sumProd / sumMag in lduMatrix).
__global__ void dotSeqK(int n, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ out)
{
    if (threadIdx.x || blockIdx.x) return;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    *out = s;
}

__global__ void sumMagSeqK(int n, const double* __restrict__ a, double* __restrict__ out)
{
    if (threadIdx.x || blockIdx.x) return;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += fabs(a[i]);
    *out = s;
}

__global__ void copyK(int n, const double* __restrict__ s, double* __restrict__ d)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = s[i];
}

} // namespace

void launchDot(Ctx& c, int n, const double* a, const double* b, double* out, int slot)
{
    if (c.sequentialReduce)
    {
        dotSeqK<<<1, 64, 0, c.stream>>>(n, a, b, out);
    }
    else
    {
        const int g = reduceGrid(n);
        dotK<<<g, kReduceBlock, 0, c.stream>>>(n, a, b, c.partials.p + (size_t)slot * kMaxReduceBlocks,
                                               c.tickets.p + slot, out);
    }
    hipCheck(hipGetLastError(), "dot");
}

void launchSumMag(Ctx& c, int n, const double* a, double* out, int slot)
{
    if (c.sequentialReduce)
    {
        sumMagSeqK<<<1, 64, 0, c.stream>>>(n, a, out);
    }
    else
    {
        const int g = reduceGrid(n);
        sumMagK<<<g, kReduceBlock, 0, c.stream>>>(n, a, c.partials.p + (size_t)slot * kMaxReduceBlocks,
                                                  c.tickets.p + slot, out);
    }
    hipCheck(hipGetLastError(), "sumMag");
}

void launchCopy(Ctx& c, int n, const double* src, double* dst)
{
    if (n <= 0) return;
    copyK<<<(n + 255) / 256, 256, 0, c.stream>>>(n, src, dst);
    hipCheck(hipGetLastError(), "copy");
}

} // namespace mfv
