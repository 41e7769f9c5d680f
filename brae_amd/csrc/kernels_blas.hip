// FP64 reductions and BLAS-1 on the device: the deterministic dot and sum|x| the PCG recurrence
// and OpenFOAM's residual norm use (reference: reductions.cu:112-147,222-250 dotKernel/sumMagKernel/
// finalSumKernel; the sequential OpenFOAM-order mode is its BRAE_DEVICE_OF_REDUCE, :36-56).
#include "reduce.hpp"

namespace mfv {

namespace {

__global__ __launch_bounds__(kReduceBlock) void dotK(int n, const double* __restrict__ a, const double* __restrict__ b,
                                                     double* __restrict__ partials)
{
    __shared__ double lds[kReduceBlock / 64];
    double v = 0.0;
    MFV_ROWS(n, i) v += a[i] * b[i];
    storePartial(blockSum(v, lds), partials);
}

__global__ __launch_bounds__(kReduceBlock) void sumMagK(int n, const double* __restrict__ a,
                                                        double* __restrict__ partials)
{
    __shared__ double lds[kReduceBlock / 64];
    double v = 0.0;
    MFV_ROWS(n, i) v += fabs(a[i]);
    storePartial(blockSum(v, lds), partials);
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
        double* pa = c.partials.p + (size_t)slot * kMaxReduceBlocks;
        dotK<<<g, kReduceBlock, 0, c.stream>>>(n, a, b, pa);
        finishK<<<1, kFinishBlock, 0, c.stream>>>(g, pa, StoreTo{out}, nullptr);
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
        double* pa = c.partials.p + (size_t)slot * kMaxReduceBlocks;
        sumMagK<<<g, kReduceBlock, 0, c.stream>>>(n, a, pa);
        finishK<<<1, kFinishBlock, 0, c.stream>>>(g, pa, StoreTo{out}, nullptr);
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
