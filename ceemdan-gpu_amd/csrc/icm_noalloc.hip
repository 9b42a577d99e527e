// One sift pass without any device allocation (icm_sift_pass_static).
//
// icm_sift_pass builds a one-pass plan per call (hipMalloc of its tables and
// knot buffers).  Callers that may not allocate - tight serving loops,
// harnesses that forbid it - use this variant instead: the same kernels, the
// same result bit for bit, with every buffer beyond the caller's input and
// output carved from one file-scope device array `scratch` of kScratchBytes
// (each GPU the library runs on holds its own copy until the process ends).
// The series tables of the I equal-length series are written on the device
// by a kernel (closed form), and every control word is cleared by a fill
// kernel, so no host-to-device copy is made either.
//
// Limits: I * N <= kMaxSamples (2^21) and I <= kMaxSeries (4096); larger
// calls return ICM_ERR_INVALID_ARG.  All calls on one GPU share that GPU's
// array and it keeps the last call's contents: two calls that overlap on one
// GPU (other threads, other streams) race and must be serialised by the
// caller.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <type_traits>

#include "../../include/iceemdan_mi355x.h"
#include "icm_errors.h"
#include "icm_kernels.h"

namespace icm {
namespace {

constexpr int64_t kMaxSamples = (int64_t)1 << 21;
constexpr int kMaxSeries = 4096;
constexpr int kTile = 256;         // samples per tile of the tile kernels
// windowed solve: halo rows per side (FP32 16, FP64 32) and rows kept per
// window, as the plans use them
template <typename real>
constexpr int solve_h() { return std::is_same<real, float>::value ? 16 : 32; }
template <typename real>
constexpr int solve_k() { return 64 * 8 - 2 * solve_h<real>(); }
constexpr size_t kScratchBytes = (size_t)24 << 20;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

// device address of `scratch` per device (hipGetSymbolAddress, once each)
std::mutex g_mu;
void* g_ptr[64] = {};

int scratch_ptr(void** out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return ICM_ERR_HIP;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return ICM_ERR_HIP;
        g_ptr[dev] = p;
    }
    *out = g_ptr[dev];
    return ICM_OK;
}

// Tables of I series of N samples each, back to back (closed form), and
// the segments of `seg` tiles, one wavefront each in the sift.
__global__ void k_uniform_series(int I, int N, int ntl, int seg, int nseg_s, int* tile_series,
                                 int* tile0, int* ntl_out, int* len, int* chan, int* nidx,
                                 int64_t* off, int64_t* kbase, int64_t* eoff, Segment* segs)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t ntiles = (int64_t)I * ntl;
    if (i < ntiles) tile_series[i] = (int)(i / ntl);
    if (i < I) {
        tile0[i] = (int)(i * ntl);
        ntl_out[i] = ntl;
        len[i] = N;
        chan[i] = 0;
        nidx[i] = (int)i;
        off[i] = i * N;
        kbase[i] = i * (int64_t)(N / 2 + 2);
        eoff[i] = i * N;
    }
    if (i < (int64_t)I * nseg_s) {
        Segment g;
        g.s = (int)(i / nseg_s);
        g.t_first = (int)(i % nseg_s) * seg;
        g.count = std::min(seg, ntl - g.t_first);
        g.pad = 0;
        segs[i] = g;
    }
}

// bump allocator over the scratch array (256-byte aligned pieces)
struct Carve {
    char* base;
    size_t used = 0;
    template <typename T>
    T* take(int64_t n)
    {
        used = (used + 255) & ~(size_t)255;
        T* p = (T*)(base + used);
        used += (size_t)std::max<int64_t>(n, 1) * sizeof(T);
        return p;
    }
};

template <typename real>
int sift_static(const real* h, int64_t N64, int64_t I64, real* h_new, int32_t* d_nk,
                hipStream_t st)
{
    const int N = (int)N64, I = (int)I64;
    const int ntl = (N + kTile - 1) / kTile;
    const int ntiles = I * ntl;
    const int seg = std::max(1, std::min(32, ntiles / 56000));
    const int nseg_s = (ntl + seg - 1) / seg;
    const int64_t kcap = N / 2 + 2;
    const int64_t kstride = (int64_t)I * kcap;
    const int64_t maxwin = (int64_t)I * 2 * ((kcap + solve_k<real>() - 1) / solve_k<real>());

    void* base = nullptr;
    int rc = scratch_ptr(&base);
    if (rc) return rc;
    Carve cv{(char*)base};
    int* tile_series = cv.take<int>(ntiles);
    int* tile0 = cv.take<int>(I);
    int* ntlv = cv.take<int>(I);
    int* len = cv.take<int>(I);
    int* chan = cv.take<int>(I);
    int* nidx = cv.take<int>(I);
    int64_t* off = cv.take<int64_t>(I);
    int64_t* kbase = cv.take<int64_t>(I);
    int64_t* eoff = cv.take<int64_t>(I);
    Segment* segs = cv.take<Segment>((int64_t)I * nseg_s);
    KnotsDev<real> k{};
    k.cnt = cv.take<int>(2 * (int64_t)ntiles);
    k.pre = cv.take<int>(2 * (int64_t)ntiles);
    k.n = cv.take<int>(2 * (int64_t)I);
    k.fix = cv.take<real>(4 * (int64_t)I);
    k.nbx = cv.take<int>(6 * (int64_t)ntiles);
    k.nby = cv.take<real>(6 * (int64_t)ntiles);
    k.mask = cv.take<uint64_t>(8 * (int64_t)ntiles);
    real* c = cv.take<real>(2 * kstride);
    int* winlist = cv.take<int>(kWinItem * maxwin);
    // control words: [0..3] window counts, [4..7] over-full counts,
    // [12] error word (epochs 1 and 2 are used)
    unsigned* ctlw = cv.take<unsigned>(16);
    if (cv.used > kScratchBytes) return ICM_ERR_INVALID_ARG;

    hipError_t e;
    const int64_t items = std::max<int64_t>(ntiles, (int64_t)I * nseg_s);
    hipLaunchKernelGGL(k_uniform_series, dim3((unsigned)((items + 255) / 256)), dim3(256), 0, st,
                       I, N, ntl, seg, nseg_s, tile_series, tile0, ntlv, len, chan, nidx, off,
                       kbase, eoff, segs);
    if ((e = hipGetLastError()) != hipSuccess) return ICM_ERR_HIP;
    if ((e = launch_fill_u32(ctlw, 0u, 16, st)) != hipSuccess) return ICM_ERR_HIP;

    SeriesDev sd{};
    sd.tile_series = tile_series;
    sd.tile0 = tile0;
    sd.ntl = ntlv;
    sd.off = off;
    sd.len = len;
    sd.kbase = kbase;
    sd.chan = chan;
    sd.eoff = eoff;
    sd.nidx = nidx;
    sd.nseries = I;
    sd.ntiles = ntiles;
    sd.max_tiles = ntl;
    sd.kstride = kstride;
    sd.cap = kTile / 2 + 2;
    int* err = (int*)(ctlw + 12);
    auto ctl = [&](uint32_t epoch) {
        PassCtl pc{};
        pc.epoch = epoch;
        pc.winlist = winlist;
        pc.wincount = ctlw;
        pc.K = solve_k<real>();
        pc.H = solve_h<real>();
        pc.err = err;
        pc.dense = ctlw + 4;
        pc.run = nullptr;
        return pc;
    };
    // knots of the input (no copy of its samples), prefixes, windows
    InitArgs<real> ia{};
    ia.mode = INIT_KNOTS;
    ia.src = h;
    if (launch_init<real>(sd, ia, nullptr, k, ctl(1), st) != hipSuccess) return ICM_ERR_HIP;
    if (launch_knot_prefix<real>(sd, k, h, ctl(1), nullptr, st) != hipSuccess) return ICM_ERR_HIP;
    SolveArgs<real> sa{};
    sa.sd = sd;
    sa.k = k;
    sa.h = h;
    sa.c = c;
    sa.winlist = winlist;
    sa.wincount = ctlw + 1;
    sa.err = err;
    if (launch_solve<real>(sa, 4096, st) != hipSuccess) return ICM_ERR_HIP;
    SiftArgs<real> a{};
    a.sd = sd;
    a.segs = segs;
    a.h_in = h;
    a.h_out = h_new;
    a.kin = k;
    a.c = c;
    a.kout = k;
    a.emit = 0;
    a.pc = ctl(2);
    a.queue = nullptr;   // one wavefront per segment (the grid is small here)
    if (launch_sift<real>(a, I * nseg_s, st) != hipSuccess) return ICM_ERR_HIP;
    if (d_nk && launch_copy<float>((float*)d_nk, (const float*)k.n, 2 * (int64_t)I, st) != hipSuccess)
        return ICM_ERR_HIP;
    int h_err = 0;
    if (hipMemcpyAsync(&h_err, err, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return ICM_ERR_HIP;
    return h_err ? ICM_ERR_DEVICE : ICM_OK;
}

}  // namespace
}  // namespace icm

extern "C" int icm_sift_pass_static(int dtype, const void* d_h, int64_t N, int64_t I,
                                    void* d_h_new, int32_t* d_nk, void* stream)
{
    if (!d_h || !d_h_new || N < 4 || N > 16000000 || I < 1 || I > icm::kMaxSeries ||
        N * I > icm::kMaxSamples)
        return ICM_ERR_INVALID_ARG;
    if (dtype == ICM_DTYPE_F64)
        return icm::sift_static<double>((const double*)d_h, N, I, (double*)d_h_new, d_nk,
                                        (hipStream_t)stream);
    if (dtype == ICM_DTYPE_F32)
        return icm::sift_static<float>((const float*)d_h, N, I, (float*)d_h_new, d_nk,
                                       (hipStream_t)stream);
    return ICM_ERR_INVALID_ARG;
}
