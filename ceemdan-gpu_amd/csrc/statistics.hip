// Statistics kernels with the reference's names and signatures
// (statistics.cuh:25-44, statistics.cu:50-240), wave64-correct and race-free:
// every thread of a block takes part in the block reduction (no early
// return before it), and one thread adds the block sum to the channel
// accumulator.
#include "../../include/statistics.h"

__global__ void multiplyS(float* input, float* output, float alpha, int N)
{
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < N) output[tid] = input[tid] * alpha;
}

__global__ void multiplyD(double* input, double* output, double alpha, int N)
{
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < N) output[tid] = input[tid] * alpha;
}

template <typename T>
__device__ void mean_body(const T* in, int N, int channels, T* chan_means)
{
    const int ch = blockIdx.y;
    if (ch >= channels) return;   // uniform over the block
    T sum = (T)0;
    const long off = (long)ch * N;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        sum += in[off + i];
    T b;
    if constexpr (sizeof(T) == 4) b = blockReduceSumS(sum);
    else b = blockReduceSumD(sum);
    if (threadIdx.x == 0) atomicAdd(&chan_means[ch], b);
}

template <typename T>
__device__ void var_body(const T* in, int N, int channels, const T* chan_means, T* chan_vars)
{
    const int ch = blockIdx.y;
    if (ch >= channels) return;
    const T mu = chan_means[ch];
    T acc = (T)0;
    const long off = (long)ch * N;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const T d = in[off + i] - mu;
        acc += d * d;
    }
    T b;
    if constexpr (sizeof(T) == 4) b = blockReduceSumS(acc);
    else b = blockReduceSumD(acc);
    if (threadIdx.x == 0) atomicAdd(&chan_vars[ch], b);
}

__global__ void meanS(float* in, const int N, const int channels, float* chan_means)
{
    mean_body<float>(in, N, channels, chan_means);
}
__global__ void meanD(double* in, const int N, const int channels, double* chan_means)
{
    mean_body<double>(in, N, channels, chan_means);
}
__global__ void varianceS(float* in, const int N, const int channels, const float* chan_means,
                          float* chan_vars)
{
    var_body<float>(in, N, channels, chan_means, chan_vars);
}
__global__ void varianceD(double* in, const int N, const int channels, const double* chan_means,
                          double* chan_vars)
{
    var_body<double>(in, N, channels, chan_means, chan_vars);
}

// Host-side launch wrappers for callers that do not compile HIP themselves
// (ctypes, other languages): same semantics as the kernels above.
#include <stdint.h>
extern "C" {
int icm_stats_mean(int dtype, void* d_in, int N, int channels, void* d_chan_means, void* stream)
{
    if (!d_in || !d_chan_means || N < 1 || channels < 1) return -1;
    const dim3 g((N + 255) / 256 < 1024 ? (N + 255) / 256 : 1024, channels);
    if (dtype == 1)
        hipLaunchKernelGGL(meanD, g, dim3(256), 0, (hipStream_t)stream, (double*)d_in, N, channels,
                           (double*)d_chan_means);
    else
        hipLaunchKernelGGL(meanS, g, dim3(256), 0, (hipStream_t)stream, (float*)d_in, N, channels,
                           (float*)d_chan_means);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}
int icm_stats_variance(int dtype, void* d_in, int N, int channels, const void* d_chan_means,
                       void* d_chan_vars, void* stream)
{
    if (!d_in || !d_chan_means || !d_chan_vars || N < 1 || channels < 1) return -1;
    const dim3 g((N + 255) / 256 < 1024 ? (N + 255) / 256 : 1024, channels);
    if (dtype == 1)
        hipLaunchKernelGGL(varianceD, g, dim3(256), 0, (hipStream_t)stream, (double*)d_in, N,
                           channels, (const double*)d_chan_means, (double*)d_chan_vars);
    else
        hipLaunchKernelGGL(varianceS, g, dim3(256), 0, (hipStream_t)stream, (float*)d_in, N,
                           channels, (const float*)d_chan_means, (float*)d_chan_vars);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}
int icm_stats_multiply(int dtype, void* d_in, void* d_out, double alpha, int N, void* stream)
{
    if (!d_in || !d_out || N < 1) return -1;
    const dim3 g((N + 255) / 256);
    if (dtype == 1)
        hipLaunchKernelGGL(multiplyD, g, dim3(256), 0, (hipStream_t)stream, (double*)d_in,
                           (double*)d_out, alpha, N);
    else
        hipLaunchKernelGGL(multiplyS, g, dim3(256), 0, (hipStream_t)stream, (float*)d_in,
                           (float*)d_out, (float)alpha, N);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}
}
