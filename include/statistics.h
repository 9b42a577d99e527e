/*
 * Statistics kernels of the MI355X ICEEMDAN library.
 *
 * Same names and signatures as the reference header statistics.cuh:25-44
 * (definitions statistics.cu:25-240), rewritten for 64-wide wavefronts: the
 * reference hard-codes 32-lane warps (statistics.cu:27,32) and writes the
 * block sum from every thread (statistics.cu:77).  Contract kept from the
 * reference: mean{S,D} ADDS the per-channel sums of in[ch*N + i] to
 * chan_means[ch] and variance{S,D} adds sum (x - chan_means[ch])^2 to
 * chan_vars[ch]; the caller zeroes the accumulators and scales with
 * multiply{S,D}.  Grid: (any x, channels); block: any multiple of 64 up to
 * 1024; no dynamic LDS is needed.  The per-block partial sums are combined
 * with atomics, so their order (and the last bits of the sum) may vary
 * between runs, as in the reference; the decomposition itself uses the
 * library's deterministic reductions instead.
 */
#ifndef ICM_STATISTICS_H
#define ICM_STATISTICS_H
#include <hip/hip_runtime.h>

/* wave64 / workgroup sums, usable from the caller's own kernels */
__device__ inline float warpReduceSumS(float val)
{
    for (int offset = 32; offset > 0; offset >>= 1) val += __shfl_xor(val, offset, 64);
    return val;
}
__device__ inline double warpReduceSumD(double val)
{
    for (int offset = 32; offset > 0; offset >>= 1) val += __shfl_xor(val, offset, 64);
    return val;
}
/* Result valid in thread 0.  Every thread of the block must call it. */
__device__ inline float blockReduceSumS(float val)
{
    __shared__ float part[16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
    val = warpReduceSumS(val);
    if (lane == 0) part[wid] = val;
    __syncthreads();
    float r = 0.0f;
    if (threadIdx.x == 0)
        for (int w = 0; w < nw; ++w) r += part[w];
    __syncthreads();
    return r;
}
__device__ inline double blockReduceSumD(double val)
{
    __shared__ double part[16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
    val = warpReduceSumD(val);
    if (lane == 0) part[wid] = val;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < nw; ++w) r += part[w];
    __syncthreads();
    return r;
}

__global__ void multiplyS(float* input, float* output, float alpha, int N);
__global__ void meanS(float* in, const int N, const int channels, float* chan_means);
__global__ void varianceS(float* in, const int N, const int channels, const float* chan_means,
                          float* chan_vars);
__global__ void multiplyD(double* input, double* output, double alpha, int N);
__global__ void meanD(double* in, const int N, const int channels, double* chan_means);
__global__ void varianceD(double* in, const int N, const int channels, const double* chan_means,
                          double* chan_vars);

#endif
