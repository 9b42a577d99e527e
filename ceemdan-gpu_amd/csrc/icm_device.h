// Device-side helpers shared by the ICEEMDAN kernels (gfx950, wave64).
//
// Everything here is written for 64-wide wavefronts: ballots are 64-bit,
// shuffles span the whole wave, and workgroup reductions combine per-wave
// partials through LDS.  No FMA contraction is allowed in this library
// (built with -ffp-contract=off): the FP64 results must round exactly like
// the CPU oracle.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace icm {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Index of this wavefront in its workgroup as a scalar value.  The compiler
// cannot prove threadIdx.x >> 6 uniform over the wave; without this, every
// per-wave quantity derived from it (work item, series, base pointers,
// lengths) would be loaded per lane and kept in vector registers.
__device__ __forceinline__ int wave_index()
{
    return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

// number of set bits of `mask` below this lane
__device__ __forceinline__ int lanes_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __ballot(p); }

template <typename T>
__device__ __forceinline__ T shfl_up(T v, int d) { return __shfl_up(v, (unsigned)d, kWave); }
template <typename T>
__device__ __forceinline__ T shfl_down(T v, int d) { return __shfl_down(v, (unsigned)d, kWave); }
template <typename T>
__device__ __forceinline__ T shfl(T v, int src) { return __shfl(v, src, kWave); }

// Whole-wave shifts by one lane (DPP wave_shr:1 / wave_shl:1): lane i gets
// lane i-1 (shift_prev) or lane i+1 (shift_next); the lane with no source
// gets `fill`.  Bit moves only, so they are exact for any type.
__device__ __forceinline__ int dpp_prev_i(int v, int fill)
{
    return __builtin_amdgcn_update_dpp(fill, v, 0x138, 0xF, 0xF, false);
}
__device__ __forceinline__ int dpp_next_i(int v, int fill)
{
    return __builtin_amdgcn_update_dpp(fill, v, 0x130, 0xF, 0xF, false);
}
__device__ __forceinline__ float shift_prev(float v, float fill)
{
    return __int_as_float(dpp_prev_i(__float_as_int(v), __float_as_int(fill)));
}
__device__ __forceinline__ float shift_next(float v, float fill)
{
    return __int_as_float(dpp_next_i(__float_as_int(v), __float_as_int(fill)));
}
__device__ __forceinline__ double shift_prev(double v, double fill)
{
    const long long b = __double_as_longlong(v), f = __double_as_longlong(fill);
    const int lo = dpp_prev_i((int)b, (int)f);
    const int hi = dpp_prev_i((int)(b >> 32), (int)(f >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double shift_next(double v, double fill)
{
    const long long b = __double_as_longlong(v), f = __double_as_longlong(fill);
    const int lo = dpp_next_i((int)b, (int)f);
    const int hi = dpp_next_i((int)(b >> 32), (int)(f >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Left and right neighbours of the samples of row v of a tile, where row p
// (q) is the row before (after) it: lane i gets lane i-1 (i+1) of v, and
// lane 0 (63) gets lane 63 (0) of p (q).  The other row's lane comes from a
// whole-wave rotation by one lane (DPP wave_ror:1 / wave_rol:1), which is
// then the fill of the shift: two moves, where a shift with a wave-uniform
// fill takes a lane read, a move and the shift.
__device__ __forceinline__ int dpp_rot_prev_i(int v) { return __builtin_amdgcn_mov_dpp(v, 0x13C, 0xF, 0xF, false); }
__device__ __forceinline__ int dpp_rot_next_i(int v) { return __builtin_amdgcn_mov_dpp(v, 0x134, 0xF, 0xF, false); }
__device__ __forceinline__ float row_prev(float v, float p)
{
    return __int_as_float(dpp_prev_i(__float_as_int(v), dpp_rot_prev_i(__float_as_int(p))));
}
__device__ __forceinline__ float row_next(float v, float q)
{
    return __int_as_float(dpp_next_i(__float_as_int(v), dpp_rot_next_i(__float_as_int(q))));
}
__device__ __forceinline__ double row_prev(double v, double p)
{
    const long long b = __double_as_longlong(v), f = __double_as_longlong(p);
    const int lo = dpp_prev_i((int)b, dpp_rot_prev_i((int)f));
    const int hi = dpp_prev_i((int)(b >> 32), dpp_rot_prev_i((int)(f >> 32)));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double row_next(double v, double q)
{
    const long long b = __double_as_longlong(v), f = __double_as_longlong(q);
    const int lo = dpp_next_i((int)b, dpp_rot_next_i((int)f));
    const int hi = dpp_next_i((int)(b >> 32), dpp_rot_next_i((int)(f >> 32)));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// DPP row moves (GFX9 encodings): row_shr:n = 0x110 + n shifts inside each
// 16-lane row, row_bcast:15 = 0x142 sends lane 15 of a row to the next row,
// row_bcast:31 = 0x143 sends lane 31 to rows 2 and 3.  Lanes with no source
// lane, or in a row left out of ROWS, keep their own `fill`.
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_row_i(int v, int fill)
{
    return __builtin_amdgcn_update_dpp(fill, v, CTRL, ROWS, 0xF, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_row(float v, float fill)
{
    return __int_as_float(dpp_row_i<CTRL, ROWS>(__float_as_int(v), __float_as_int(fill)));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_row(double v, double fill)
{
    const long long b = __double_as_longlong(v), f = __double_as_longlong(fill);
    const int lo = dpp_row_i<CTRL, ROWS>((int)b, (int)f);
    const int hi = dpp_row_i<CTRL, ROWS>((int)(b >> 32), (int)(f >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <int CTRL, int ROWS, typename real, int NV>
__device__ __forceinline__ void seg_scan_step(real (&v)[NV], int key)
{
    // a lane that receives nothing sees key + 1, which never matches
    const bool take = dpp_row_i<CTRL, ROWS>(key, key + 1) == key;
#pragma unroll
    for (int z = 0; z < NV; ++z) {
        const real w = dpp_row<CTRL, ROWS>(v[z], (real)0);
        v[z] = take ? v[z] + w : v[z];
    }
}

// Inclusive scan of NV values per lane over the runs of equal `key`, for
// keys that do not decrease with the lane number: lane i ends with the sum
// over the lanes j <= i of its run.  Steps of 1, 2, 4, 8 lanes inside each
// row, then the last lane of row 0 (2) into row 1 (3), then lane 31 into
// rows 2 and 3; with non-decreasing keys, "the source lane has my key"
// means every lane between belongs to my run.  Every lane executes it.
template <typename real, int NV>
__device__ __forceinline__ void seg_scan_keys(real (&v)[NV], int key)
{
    seg_scan_step<0x111, 0xF>(v, key);
    seg_scan_step<0x112, 0xF>(v, key);
    seg_scan_step<0x114, 0xF>(v, key);
    seg_scan_step<0x118, 0xF>(v, key);
    seg_scan_step<0x142, 0xA>(v, key);
    seg_scan_step<0x143, 0xC>(v, key);
}

// Wave-uniform load through the scalar data cache (constant address space).
// Only for data that the running kernel does not write.
template <typename T>
__device__ __forceinline__ T load_scalar(const T* p)
{
    return *(const __attribute__((address_space(4))) T*)p;
}

// Wait until every vector memory access of this wave has completed
// (s_waitcnt vmcnt(0); expcnt and lgkmcnt left at their maxima).
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// value of lane `l` (wave-uniform result)
__device__ __forceinline__ float lane_value(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ uint64_t lane_value(uint64_t v, int l)
{
    const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)v, l);
    const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ double lane_value(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// ------------------------------------------------------------------ Philox
struct u64x4 {
    uint64_t v[4];
};

__device__ __forceinline__ u64x4 philox4x64_10(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3,
                                              uint64_t k0, uint64_t k1)
{
    const uint64_t M0 = 0xD2E7470EE14C6C93ULL, M1 = 0xCA5A826395121157ULL;
    const uint64_t W0 = 0x9E3779B97F4A7C15ULL, W1 = 0xBB67AE8584CAA73BULL;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t lo0 = M0 * c0, hi0 = __umul64hi(M0, c0);
        uint64_t lo1 = M1 * c2, hi1 = __umul64hi(M1, c2);
        uint64_t n0 = hi1 ^ c1 ^ k0;
        uint64_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += W0; k1 += W1;
    }
    u64x4 o;
    o.v[0] = c0; o.v[1] = c1; o.v[2] = c2; o.v[3] = c3;
    return o;
}

}  // namespace icm
