// HIP kernels of the ICEEMDAN hot path (gfx950, wave64).
//
// A knot set is stored as bitmasks: 64-bit words, one per 64 samples of a
// 256-sample tile and knot type, written by the kernel that wrote the
// samples; a knot's position is its bit, its value the sample there.  One
// sift pass of every series is three launches:
//   k_solve         one wavefront per window of 64*B knots: positions from
//                   the bitmasks, values from the samples, natural-spline
//                   rows, per-lane elimination of B rows, parallel cyclic
//                   reduction across the 64 lanes, back substitution -> c
//   k_sift_w        one wavefront per segment of tiles: interval of every
//                   sample from the tile's bitmasks, interval coefficients in
//                   LDS, upper/lower envelope, mean, subtract; then the
//                   extrema of the new samples as the next pass's bitmasks
//   k_knot_prefix_* per series: tile knot prefixes, neighbour knots of every
//                   tile, the solve's window list, the count of tiles too
//                   full for the small knot table
// From the second stage of a phase on, k_sift_w runs with a small knot table
// (more waves per SIMD) and k_sift_list sifts, with the full table, the tiles
// it could not hold (or every segment, when the count says they are many).
// Reference counterparts: find_extrema_shfl_* :92-148, CUB ExclusiveSum
// :859-883, select_extrema_* :150-207, setBoundary* :209-291,
// tridiagonal_setup :310-343, gtsv2_nopivot :899-921, spline_coefficients
// :382-401, interpolate :403-439, averageUppperLower :441-454,
// produceSX/thresholdJudge :456-492, updateRealizations :527-544
// (all cudaICEEMDAN.cu).
#include "icm_device.h"
#include "icm_kernels.h"

#include <map>
#include <mutex>
#include <type_traits>
#include <utility>
#include "icm_errors.h"

namespace icm {

// promotion type used where the reference mixes float data with double
// literals (cudaICEEMDAN.cu:50, :64, :76); identity for double
template <typename real> struct Promote { using type = double; };

// ------------------------------------------------------------ tile helpers

constexpr int kWT = 256;               // samples per tile
constexpr int kWR = kWT / kWave;       // 64-sample chunks per tile
constexpr int kWKC = kWT / 2 + 8;      // table entries per type (>= cnt + 3 for any tile)
constexpr int kWKS = kWave;             // small table (sparse passes): tiles with cnt + 3 <= 64

// ------------------------------------------------------------- knot bitmasks
// A tile's knots of one type are the set bits of four 64-bit row masks
// m[r] (bit b of row r = sample 64 r + b of the tile).

// tile offset of the last set bit, -1 if none
__device__ __forceinline__ int mask_last(const uint64_t (&m)[kWR])
{
    int last = -1;
#pragma unroll
    for (int r = 0; r < kWR; ++r) last = m[r] ? kWave * r + 63 - __clzll(m[r]) : last;
    return last;
}

// orders this wave's LDS and memory accesses before and after it
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Bitmasks of the strict maxima and minima among one row of 64 samples
// (v against its neighbours lft, rgt).  Each comparison is balloted on its
// own and the masks are combined by scalar instructions: the ballot of one
// comparison is that comparison's lane mask, while the ballot of a
// combination of flags costs a v_cndmask and a v_cmp (the compiler turns
// the flag into a 0/1 vector value and compares it again).  The empty asm
// statement makes the two row masks scalar registers at this point, so the
// four comparison masks of every row are not all held until the masks are
// stored (held together they spilled).  `valid` and `end` are the masks of
// the lanes inside the series and at its two ends (ends are knots of both
// types); FULL rows have every lane valid and no end, and ignore both.
template <bool FULL, typename real>
__device__ __forceinline__ void row_extrema(real v, real lft, real rgt, uint64_t valid, uint64_t end,
                                            uint64_t& mx, uint64_t& mn)
{
    const uint64_t gl = wave_ballot(v > lft), gr = wave_ballot(v > rgt);
    const uint64_t ll = wave_ballot(v < lft), lr = wave_ballot(v < rgt);
    if (FULL) {
        mx = gl & gr;
        mn = ll & lr;
    } else {
        mx = valid & (end | (gl & gr));
        mn = valid & (end | (ll & lr));
    }
    asm volatile("" : "+s"(mx), "+s"(mn));
}

// tile offset of the first set bit, -1 if none
__device__ __forceinline__ int mask_first(const uint64_t (&m)[kWR])
{
    int first = -1;
#pragma unroll
    for (int r = kWR - 1; r >= 0; --r) first = m[r] ? kWave * r + __ffsll((long long)m[r]) - 1 : first;
    return first;
}

template <typename real>
__device__ __forceinline__ void load_masks(const KnotsDev<real>& k, int64_t gt, int type,
                                           uint64_t (&m)[kWR])
{
#pragma unroll
    for (int r = 0; r < kWR; ++r) m[r] = k.mask[8 * gt + kWR * type + r];
}

struct LastTwo {
    int c;
    int64_t l1, l2;
};

__device__ __forceinline__ LastTwo lt_combine(const LastTwo& A, const LastTwo& B)
{
    // written as selects (no early returns) so the result stays in registers
    LastTwo r;
    r.c = B.c >= 2 ? 2 : (B.c == 1 ? min(A.c + 1, 2) : A.c);
    r.l1 = B.c >= 1 ? B.l1 : A.l1;
    r.l2 = B.c >= 2 ? B.l2 : (B.c == 1 ? A.l1 : A.l2);
    return r;
}

__device__ __forceinline__ LastTwo lt_shfl_up(const LastTwo& v, int o)
{
    LastTwo r;
    r.c = shfl_up(v.c, o);
    r.l1 = shfl_up(v.l1, o);
    r.l2 = shfl_up(v.l2, o);
    return r;
}

// Knot prefix: per-series exclusive scan of the tile knot counts (the CUB
// ExclusiveSum of cudaICEEMDAN.cu:859-883, over tiles instead of samples),
// the series knot counts, the solve's window items, and for every tile the
// positions and values of the two knots before it and of the first knot
// after it (positions from the bitmasks, values from the pass input h).
//
// The same outputs with one 1024-thread workgroup per series and one tile per
// thread (chunks of 1024 tiles): workgroup scans forward (counts, last two
// knots) and backward (next knot), carried from chunk to chunk.
constexpr int kPT = 1024;

__device__ __forceinline__ LastTwo lt_none()
{
    LastTwo r;
    r.c = 0;
    r.l1 = -1;
    r.l2 = -1;
    return r;
}

// Window item w of a series with wmx windows of maxima and wmn of minima:
// the two types alternate (maxima window i, then minima window i), so the
// two windows over one stretch of samples go to neighbouring waves of one
// solve workgroup and read the same cache lines at about the same time.
__device__ __forceinline__ void window_of_item(int w, int wmx, int wmn, int& type, int& wi)
{
    const int m = min(wmx, wmn);
    if (w < 2 * m) {
        type = w & 1;
        wi = w >> 1;
    } else {
        type = wmx > wmn ? 0 : 1;
        wi = m + (w - 2 * m);
    }
}

template <typename real>
__global__ __launch_bounds__(kPT) void k_knot_prefix_b(SeriesDev sd, KnotsDev<real> k, const real* h,
                                                       PassCtl pc, const int* active)
{
    constexpr int NWV = kPT / kWave;
    constexpr int INF = 0x7FFFFFFF;
    __shared__ int s_tot[NWV], s_ex_tot[NWV], s_all;
    __shared__ LastTwo s_lt[NWV], s_ex_lt[NWV], s_all_lt;
    __shared__ int s_min[NWV];
    __shared__ unsigned s_wbase;
    const int s = blockIdx.x;
    if (pc.run && load_scalar(pc.run) == 0u) return;
    if (active && !active[s]) return;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int tb = sd.tile0[s], nt = sd.ntl[s];
    const real* hs = h + sd.off[s];
    {
        // tiles the small sift table cannot hold
        unsigned d = 0;
        for (int t = tid; t < nt; t += kPT) {
            const int c0 = k.cnt[2 * (tb + t)], c1 = k.cnt[2 * (tb + t) + 1];
            d += (c0 + 3 > kWKS || c1 + 3 > kWKS) ? 1u : 0u;
        }
        for (int o = 32; o >= 1; o >>= 1) d += __shfl_xor(d, o, kWave);
        if (lane == 0 && d) atomicAdd(&pc.dense[pc.epoch], d);
    }
    int ntot[2] = {0, 0};
    for (int ty = 0; ty < 2; ++ty) {
        // forward: exclusive count and the last two knots before each tile
        int carry = 0;
        LastTwo carry_lt = lt_none();
        for (int base = 0; base < nt; base += kPT) {
            const int t = base + tid;
            const int c = t < nt ? k.cnt[2 * (tb + t) + ty] : 0;
            LastTwo own = lt_none();
            if (c > 0) {
                uint64_t m[kWR];
                load_masks(k, tb + t, ty, m);
                const int b1 = mask_last(m);
                m[b1 / kWave] ^= 1ull << (b1 % kWave);
                own.c = min(c, 2);
                own.l1 = (int64_t)t * kWT + b1;
                own.l2 = c >= 2 ? (int64_t)t * kWT + mask_last(m) : -1;
            }
            int inc = c;
            LastTwo sc = own;
            for (int o = 1; o < kWave; o <<= 1) {
                const int v = shfl_up(inc, o);
                const LastTwo w = lt_shfl_up(sc, o);
                if (lane >= o) {
                    inc += v;
                    sc = lt_combine(w, sc);
                }
            }
            const int xi = shfl_up(inc, 1);
            const LastTwo xl = lt_shfl_up(sc, 1);
            __syncthreads();   // earlier readers of s_* are done
            if (lane == kWave - 1) {
                s_tot[wv] = inc;
                s_lt[wv] = sc;
            }
            __syncthreads();
            if (wv == 0) {
                // exclusive scan of the wave totals by the first NWV lanes
                int v = lane < NWV ? s_tot[lane] : 0;
                LastTwo l = lane < NWV ? s_lt[lane] : lt_none();
                int vi = v;
                LastTwo li = l;
                for (int o = 1; o < NWV; o <<= 1) {
                    const int a1 = shfl_up(vi, o);
                    const LastTwo b1 = lt_shfl_up(li, o);
                    if (lane >= o) {
                        vi += a1;
                        li = lt_combine(b1, li);
                    }
                }
                const LastTwo le = lt_shfl_up(li, 1);
                if (lane < NWV) {
                    s_ex_tot[lane] = vi - v;
                    s_ex_lt[lane] = lane > 0 ? le : lt_none();
                }
                if (lane == NWV - 1) {
                    s_all = vi;
                    s_all_lt = li;
                }
            }
            __syncthreads();
            const int ex = carry + s_ex_tot[wv] + (lane > 0 ? xi : 0);
            const LastTwo before =
                lt_combine(carry_lt, lt_combine(s_ex_lt[wv], lane > 0 ? xl : lt_none()));
            if (t < nt) {
                k.pre[2 * (tb + t) + ty] = ex;
                const int64_t o = 6 * (int64_t)(tb + t) + 3 * ty;
                k.nbx[o] = before.l1 >= 0 ? (int)before.l1 : -1;
                k.nby[o] = before.l1 >= 0 ? hs[before.l1] : (real)0;
                k.nbx[o + 1] = before.l2 >= 0 ? (int)before.l2 : -1;
                k.nby[o + 1] = before.l2 >= 0 ? hs[before.l2] : (real)0;
            }
            carry += s_all;
            carry_lt = lt_combine(carry_lt, s_all_lt);
        }
        ntot[ty] = carry;
        // backward: first non-empty tile after each tile
        int carry_min = INF;
        for (int base = ((nt - 1) / kPT) * kPT; base >= 0; base -= kPT) {
            const int t = base + tid;
            const int c = t < nt ? k.cnt[2 * (tb + t) + ty] : 0;
            int mn = c > 0 ? t : INF;
            for (int o = 1; o < kWave; o <<= 1) {
                const int m = shfl_down(mn, o);
                if (lane + o < kWave) mn = min(mn, m);
            }
            const int xm = shfl_down(mn, 1);
            __syncthreads();
            if (lane == 0) s_min[wv] = mn;
            __syncthreads();
            int after = min(carry_min, lane < kWave - 1 ? xm : INF);
            int all = carry_min;
            for (int w = 0; w < NWV; ++w) {
                if (w > wv) after = min(after, s_min[w]);
                all = min(all, s_min[w]);
            }
            if (t < nt) {
                const int64_t o = 6 * (int64_t)(tb + t) + 3 * ty + 2;
                int x3 = -1;
                if (after != INF) {
                    uint64_t m[kWR];
                    load_masks(k, tb + after, ty, m);
                    x3 = after * kWT + mask_first(m);
                }
                k.nbx[o] = x3;
                k.nby[o] = x3 >= 0 ? hs[x3] : (real)0;
            }
            carry_min = all;
        }
    }
    const int wmx = (ntot[0] + pc.K - 1) / pc.K;
    const int wmn = (ntot[1] + pc.K - 1) / pc.K;
    if (tid == 0) {
        k.n[2 * s] = ntot[0];
        k.n[2 * s + 1] = ntot[1];
        s_wbase = atomicAdd(&pc.wincount[pc.epoch], (unsigned)(wmx + wmn));
    }
    __syncthreads();
    const unsigned wbase = s_wbase;
    for (int w = tid; w < wmx + wmn; w += kPT) {
        int type, wi;
        window_of_item(w, wmx, wmn, type, wi);
        pc.winlist[kWinItem * (wbase + w)] = s;
        pc.winlist[kWinItem * (wbase + w) + 1] = (wi << 1) | type;
        pc.winlist[kWinItem * (wbase + w) + 2] = -1;   // no tile hint: the solve searches
    }
}

// Series of at most 1024 tiles: the same outputs in one pass, both knot
// types and both scan directions at once (one load round, two workgroup
// barriers, one gather round).
template <typename real>
__global__ __launch_bounds__(kPT) void k_knot_prefix_1(SeriesDev sd, KnotsDev<real> k, const real* h,
                                                       PassCtl pc, const int* active)
{
    constexpr int NWV = kPT / kWave;
    constexpr int INF = 0x7FFFFFFF;
    __shared__ int s_tot[2][NWV], s_ex_tot[2][NWV], s_all[2];
    __shared__ LastTwo s_lt[2][NWV], s_ex_lt[2][NWV];
    __shared__ int s_min[2][NWV], s_ex_min[2][NWV];
    __shared__ int s_pre[2][kPT];
    __shared__ int s_first[2][kPT];
    __shared__ unsigned s_wbase;
    __shared__ unsigned s_dense[NWV];
    const int s = blockIdx.x;
    if (pc.run && load_scalar(pc.run) == 0u) return;
    if (active && !active[s]) return;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int tb = sd.tile0[s], nt = sd.ntl[s];
    const real* hs = h + sd.off[s];
    const int t = tid;
    int c[2];
    LastTwo sc[2];
    int inc[2], mn[2];
    int first[2];         // tile offset of the tile's first knot per type
    // both types' counts, then both types' bitmasks, in one round trip each
    c[0] = t < nt ? k.cnt[2 * (tb + t)] : 0;
    c[1] = t < nt ? k.cnt[2 * (tb + t) + 1] : 0;
    // the bitmasks loaded with the counts (not after them, for the tiles
    // with knots only): one round trip fewer per workgroup
    uint64_t mm[2][kWR];
#pragma unroll
    for (int q = 0; q < 2 * kWR; ++q)
        mm[q / kWR][q % kWR] = t < nt ? k.mask[8 * (int64_t)(tb + min(t, nt - 1)) + q] : 0ull;
#pragma unroll
    for (int ty = 0; ty < 2; ++ty) {
        sc[ty] = lt_none();
        first[ty] = -1;
        if (c[ty] > 0) {
            uint64_t m[kWR];
#pragma unroll
            for (int r = 0; r < kWR; ++r) m[r] = mm[ty][r];
            first[ty] = mask_first(m);
            const int b1 = mask_last(m);
            m[b1 / kWave] ^= 1ull << (b1 % kWave);
            sc[ty].c = min(c[ty], 2);
            sc[ty].l1 = (int64_t)t * kWT + b1;
            sc[ty].l2 = c[ty] >= 2 ? (int64_t)t * kWT + mask_last(m) : -1;
        }
        inc[ty] = c[ty];
        mn[ty] = c[ty] > 0 ? t : INF;
    }
    {
        // tiles the small sift table cannot hold
        const unsigned d = (unsigned)__popcll(wave_ballot(c[0] + 3 > kWKS || c[1] + 3 > kWKS));
        if (lane == 0) s_dense[wv] = d;
    }
    for (int o = 1; o < kWave; o <<= 1) {
#pragma unroll
        for (int ty = 0; ty < 2; ++ty) {
            const int v = shfl_up(inc[ty], o);
            const LastTwo w = lt_shfl_up(sc[ty], o);
            const int m = shfl_down(mn[ty], o);
            if (lane >= o) {
                inc[ty] += v;
                sc[ty] = lt_combine(w, sc[ty]);
            }
            if (lane + o < kWave) mn[ty] = min(mn[ty], m);
        }
    }
    int xi[2], xm[2];
    LastTwo xl[2];
#pragma unroll
    for (int ty = 0; ty < 2; ++ty) {
        xi[ty] = shfl_up(inc[ty], 1);
        xl[ty] = lt_shfl_up(sc[ty], 1);
        xm[ty] = shfl_down(mn[ty], 1);
        if (lane == kWave - 1) {
            s_tot[ty][wv] = inc[ty];
            s_lt[ty][wv] = sc[ty];
        }
        if (lane == 0) s_min[ty][wv] = mn[ty];
        s_first[ty][t] = first[ty];
    }
    __syncthreads();
    if (wv == 0) {
        // scans of the wave totals by the first NWV lanes
#pragma unroll
        for (int ty = 0; ty < 2; ++ty) {
            const int v = lane < NWV ? s_tot[ty][lane] : 0;
            const LastTwo l = lane < NWV ? s_lt[ty][lane] : lt_none();
            const int m0 = lane < NWV ? s_min[ty][lane] : INF;
            int vi = v, mi = m0;
            LastTwo li = l;
            for (int o = 1; o < NWV; o <<= 1) {
                const int a1 = shfl_up(vi, o);
                const LastTwo b1 = lt_shfl_up(li, o);
                const int c1 = shfl_down(mi, o);
                if (lane >= o) {
                    vi += a1;
                    li = lt_combine(b1, li);
                }
                if (lane + o < NWV) mi = min(mi, c1);
            }
            const LastTwo le = lt_shfl_up(li, 1);
            const int me = shfl_down(mi, 1);
            if (lane < NWV) {
                s_ex_tot[ty][lane] = vi - v;
                s_ex_lt[ty][lane] = lane > 0 ? le : lt_none();
                s_ex_min[ty][lane] = lane < NWV - 1 ? me : INF;
            }
            if (lane == NWV - 1) s_all[ty] = vi;
        }
    }
    __syncthreads();
    if (t < nt) {
        // positions of both types' neighbour knots, then all six value loads
        // (at clamped positions, so that none sits under a condition), then
        // the stores
        int ex[2], xs[2][3];
#pragma unroll
        for (int ty = 0; ty < 2; ++ty) {
            ex[ty] = s_ex_tot[ty][wv] + (lane > 0 ? xi[ty] : 0);
            const LastTwo before = lt_combine(s_ex_lt[ty][wv], lane > 0 ? xl[ty] : lt_none());
            const int after = min(s_ex_min[ty][wv], lane < kWave - 1 ? xm[ty] : INF);
            xs[ty][0] = before.l1 >= 0 ? (int)before.l1 : -1;
            xs[ty][1] = before.l2 >= 0 ? (int)before.l2 : -1;
            xs[ty][2] = after == INF ? -1 : after * kWT + s_first[ty][after];
        }
        real ys[2][3];
#pragma unroll
        for (int ty = 0; ty < 2; ++ty)
#pragma unroll
            for (int w = 0; w < 3; ++w) ys[ty][w] = hs[max(xs[ty][w], 0)];
        if (tid == 0) {
            // the window base is requested while the values are in flight
            const int n0 = s_all[0], n1 = s_all[1];
            k.n[2 * s] = n0;
            k.n[2 * s + 1] = n1;
            s_wbase = atomicAdd(&pc.wincount[pc.epoch],
                                (unsigned)((n0 + pc.K - 1) / pc.K + (n1 + pc.K - 1) / pc.K));
        }
#pragma unroll
        for (int ty = 0; ty < 2; ++ty) {
            k.pre[2 * (tb + t) + ty] = ex[ty];
            s_pre[ty][t] = ex[ty];
            const int64_t o = 6 * (int64_t)(tb + t) + 3 * ty;
#pragma unroll
            for (int w = 0; w < 3; ++w) {
                k.nbx[o + w] = xs[ty][w];
                k.nby[o + w] = xs[ty][w] >= 0 ? ys[ty][w] : (real)0;
            }
        }
    }
    const int n0 = s_all[0], n1 = s_all[1];
    const int wmx = (n0 + pc.K - 1) / pc.K;
    const int wmn = (n1 + pc.K - 1) / pc.K;
    if (tid == 0 && nt == 0) {
        k.n[2 * s] = n0;
        k.n[2 * s + 1] = n1;
        s_wbase = atomicAdd(&pc.wincount[pc.epoch], (unsigned)(wmx + wmn));
    }
    if (tid == 0) {
        unsigned d = 0;
        for (int w = 0; w < NWV; ++w) d += s_dense[w];
        if (d) atomicAdd(&pc.dense[pc.epoch], d);
    }
    __syncthreads();
    const unsigned wbase = s_wbase;
    for (int w = tid; w < wmx + wmn; w += kPT) {
        int type, wi;
        window_of_item(w, wmx, wmn, type, wi);
        // the last tile whose knots start at or before the window's first
        // gathered knot, so the solve needs no search
        const int glo = max(wi * pc.K - pc.H - 1, 0);
        int x = 0, y = nt - 1;
        while (x < y) {
            const int mid = (x + y + 1) >> 1;
            if (s_pre[type][mid] <= glo) x = mid;
            else y = mid - 1;
        }
        pc.winlist[kWinItem * (wbase + w)] = s;
        pc.winlist[kWinItem * (wbase + w) + 1] = (wi << 1) | type;
        pc.winlist[kWinItem * (wbase + w) + 2] = x;
    }
}

// Series of at most kPT tiles: the same outputs with one wavefront per
// series and no workgroup barrier.  Lane l owns tile 64 j + l of chunk j;
// the wave scans each chunk (counts, and the last two knots before each
// tile) and carries the totals to the next chunk, with the next chunk's
// counts and bitmasks in flight meanwhile; a backward sweep over the chunks
// (suffix minimum of the tiles' first knots) gives the first knot after each
// tile.  Each window's first tile is known from the scan (the tile holding
// the window's first gathered knot) and kept in LDS until the series' window
// count is known.  Workgroups are four independent waves, so beside another
// stream's sift the kernel runs in whatever wave slots that kernel leaves
// free (a 1024-thread workgroup needs 16 free slots on one CU at once, and
// then waited for the other stream's sift to drain: 3.5 ms per pass beside
// it against 0.5 ms alone).
struct Two {
    int c, l1, l2;   // number of knots seen (at most 2), the last, the one before
};

__device__ __forceinline__ Two two_none() { return Two{0, -1, -1}; }

__device__ __forceinline__ Two two_combine(const Two& A, const Two& B)
{
    Two r;
    r.c = B.c >= 2 ? 2 : (B.c == 1 ? min(A.c + 1, 2) : A.c);
    r.l1 = B.c >= 1 ? B.l1 : A.l1;
    r.l2 = B.c >= 2 ? B.l2 : (B.c == 1 ? A.l1 : A.l2);
    return r;
}

__device__ __forceinline__ Two two_shfl_up(const Two& v, int o)
{
    return Two{shfl_up(v.c, o), shfl_up(v.l1, o), shfl_up(v.l2, o)};
}

__device__ __forceinline__ Two two_lane(const Two& v, int l)
{
    return Two{__builtin_amdgcn_readlane(v.c, l), __builtin_amdgcn_readlane(v.l1, l),
                __builtin_amdgcn_readlane(v.l2, l)};
}

// windows per type of a series of at most kPT tiles: <= 129 knots per tile
// and type, K >= 448 rows kept per window
constexpr int kPWin = (kPT * 129 + 447) / 448 + 1;
// series length (tiles) up to which the one-wave kernel is used
constexpr int kPWTiles = 128;

template <typename real>
__global__ __launch_bounds__(kThreads) void k_knot_prefix_w(SeriesDev sd, KnotsDev<real> k, const real* h,
                                                            PassCtl pc, const int* active)
{
    constexpr int NW = kThreads / kWave;
    constexpr int INF = 0x7FFFFFFF;
    __shared__ int s_hint[NW][2][kPWin];
    const int lane = threadIdx.x & 63, wv = wave_index();
    const int s = blockIdx.x * NW + wv;
    if (s >= sd.nseries) return;
    if (pc.run && load_scalar(pc.run) == 0u) return;
    if (active && !load_scalar(active + s)) return;
    const int tb = load_scalar(sd.tile0 + s), nt = load_scalar(sd.ntl + s);
    const real* hs = h + load_scalar(sd.off + s);
    const int K = pc.K, H = pc.H;
    const int nch = (nt + kWave - 1) / kWave;

    // counts and the eight row bitmasks of tile 64 ch + lane (zero past the series)
    auto load = [&](int ch, int (&c)[2], uint64_t (&m)[2 * kWR]) {
        const int t = ch * kWave + lane;
        const bool in = t < nt;
        const int64_t gt = tb + (in ? t : 0);
        c[0] = in ? k.cnt[2 * gt] : 0;
        c[1] = in ? k.cnt[2 * gt + 1] : 0;
#pragma unroll
        for (int q = 0; q < 2 * kWR; ++q) m[q] = in ? k.mask[8 * gt + q] : 0ull;
    };

    int carry[2] = {0, 0};
    Two carry_lt[2] = {two_none(), two_none()};
    unsigned dense = 0;   // tiles the small sift table cannot hold
    int c[2], cn[2] = {0, 0};
    uint64_t m[2 * kWR], mn[2 * kWR] = {};
    load(0, c, m);
    for (int ch = 0; ch < nch; ++ch) {
        if (ch + 1 < nch) load(ch + 1, cn, mn);
        const int t = ch * kWave + lane;
        const bool in = t < nt;
        const int64_t gt = tb + t;
        dense += (unsigned)__popcll(wave_ballot(in && (c[0] + 3 > kWKS || c[1] + 3 > kWKS)));
        int ex[2], xb[2][2], fp[2];
#pragma unroll
        for (int ty = 0; ty < 2; ++ty) {
            uint64_t mm[kWR];
#pragma unroll
            for (int r = 0; r < kWR; ++r) mm[r] = m[ty * kWR + r];
            Two own = two_none();
            fp[ty] = -1;
            if (c[ty] > 0) {
                fp[ty] = t * kWT + mask_first(mm);
                const int b1 = mask_last(mm);
                mm[b1 / kWave] ^= 1ull << (b1 % kWave);
                own.c = min(c[ty], 2);
                own.l1 = t * kWT + b1;
                own.l2 = c[ty] >= 2 ? t * kWT + mask_last(mm) : -1;
            }
            int inc = c[ty];
            Two sc = own;
#pragma unroll
            for (int o = 1; o < kWave; o <<= 1) {
                const int v = shfl_up(inc, o);
                const Two w = two_shfl_up(sc, o);
                if (lane >= o) {
                    inc += v;
                    sc = two_combine(w, sc);
                }
            }
            ex[ty] = carry[ty] + inc - c[ty];
            const Two xl = two_shfl_up(sc, 1);
            const Two before = two_combine(carry_lt[ty], lane > 0 ? xl : two_none());
            xb[ty][0] = before.l1;
            xb[ty][1] = before.l2;
            // the window whose first gathered knot (max(w K - H - 1, 0)) lies
            // in this tile: at most one, as K exceeds a tile's knots
            if (c[ty] > 0) {
                if (ex[ty] == 0) s_hint[wv][ty][0] = t;
                const int w = (ex[ty] + c[ty] + H) / K;
                if (w >= 1 && w < kPWin && w * K - H - 1 >= ex[ty]) s_hint[wv][ty][w] = t;
            }
            carry[ty] += __builtin_amdgcn_readlane(inc, kWave - 1);
            carry_lt[ty] = two_combine(carry_lt[ty], two_lane(sc, kWave - 1));
        }
        // values of the knots before the tile (positions of knots that exist,
        // or 0), all loads before the stores
        real yb[2][2];
#pragma unroll
        for (int ty = 0; ty < 2; ++ty)
#pragma unroll
            for (int w = 0; w < 2; ++w) yb[ty][w] = hs[max(xb[ty][w], 0)];
        if (in) {
#pragma unroll
            for (int ty = 0; ty < 2; ++ty) {
                k.pre[2 * gt + ty] = ex[ty];
                const int64_t o = 6 * gt + 3 * ty;
                k.nbx[o] = xb[ty][0];
                k.nby[o] = xb[ty][0] >= 0 ? yb[ty][0] : (real)0;
                k.nbx[o + 1] = xb[ty][1];
                k.nby[o + 1] = xb[ty][1] >= 0 ? yb[ty][1] : (real)0;
                // the tile's own first knot, read back by this lane below
                k.nbx[o + 2] = fp[ty];
            }
        }
        c[0] = cn[0];
        c[1] = cn[1];
#pragma unroll
        for (int q = 0; q < 2 * kWR; ++q) m[q] = mn[q];
    }
    // backward: the first knot after each tile = the least first knot of
    // the tiles after it
    int cmin[2] = {INF, INF};
    for (int ch = nch - 1; ch >= 0; --ch) {
        const int t = ch * kWave + lane;
        const bool in = t < nt;
        const int64_t gt = tb + (in ? t : 0);
        int after[2];
#pragma unroll
        for (int ty = 0; ty < 2; ++ty) {
            const int f = in ? k.nbx[6 * gt + 3 * ty + 2] : -1;
            int mnv = f >= 0 ? f : INF;
#pragma unroll
            for (int o = 1; o < kWave; o <<= 1) {
                const int v = shfl_down(mnv, o);
                if (lane + o < kWave) mnv = min(mnv, v);
            }
            const int nx = shfl_down(mnv, 1);
            after[ty] = min(cmin[ty], lane < kWave - 1 ? nx : INF);
            cmin[ty] = min(cmin[ty], __builtin_amdgcn_readlane(mnv, 0));
        }
        real ya[2];
#pragma unroll
        for (int ty = 0; ty < 2; ++ty) ya[ty] = hs[after[ty] == INF ? 0 : after[ty]];
        if (in) {
#pragma unroll
            for (int ty = 0; ty < 2; ++ty) {
                const int64_t o = 6 * gt + 3 * ty + 2;
                k.nbx[o] = after[ty] == INF ? -1 : after[ty];
                k.nby[o] = after[ty] == INF ? (real)0 : ya[ty];
            }
        }
    }
    const int n0 = carry[0], n1 = carry[1];
    const int wmx = (n0 + K - 1) / K;
    const int wmn = (n1 + K - 1) / K;
    unsigned wbase = 0;
    if (lane == 0) {
        k.n[2 * s] = n0;
        k.n[2 * s + 1] = n1;
        wbase = atomicAdd(&pc.wincount[pc.epoch], (unsigned)(wmx + wmn));
        if (dense) atomicAdd(&pc.dense[pc.epoch], dense);
    }
    wbase = (unsigned)__builtin_amdgcn_readfirstlane((int)wbase);
    wave_sync();   // the window hints written by other lanes
    for (int w = lane; w < wmx + wmn; w += kWave) {
        int type, wi;
        window_of_item(w, wmx, wmn, type, wi);
        pc.winlist[kWinItem * (wbase + w)] = s;
        pc.winlist[kWinItem * (wbase + w) + 1] = (wi << 1) | type;
        pc.winlist[kWinItem * (wbase + w) + 2] = wi < kPWin ? s_hint[wv][type][wi] : -1;
    }
}

// ------------------------------------------------------------- stage init

template <typename real>
__device__ __forceinline__ real init_value(const InitArgs<real>& ia, const SeriesDev& sd, int s,
                                           int p)
{
    using P = typename Promote<real>::type;
    if (ia.mode == INIT_COPY) return ia.src[sd.off[s] + p];
    if (ia.mode == INIT_RESID) return ia.src[sd.off[s] + p] - ia.E[sd.off[s] + p];
    const int c = sd.chan[s];
    const real base = ia.chan_sig[ia.chan_off[c] + p];
    const real e = ia.E[sd.eoff[s] + p];
    if (ia.mode == INIT_STAGE0) {
        // x_i = x~ + eps0 * E_1(w^i) * (1 / std(E_1(w^i)))   (cudaICEEMDAN.cu:76)
        return (real)((P)base + (P)(ia.eps0 * e) * (P)ia.scale[sd.nidx[s]]);
    }
    // x_i = r_k + eps0 * E_{k+1}(w^i) * std(r_k)           (cudaICEEMDAN.cu:88)
    return base + (ia.eps0 * e) * (real)ia.scale[c];
}

// ---------------------------------------------------------------- sift pass

// Interval coefficients S(t) = a + (b + (c + d t) t) t, t = p - pos.
template <typename real>
struct alignas(2 * sizeof(real)) Coef {
    real a, b, c, d;
};

// Horner steps as fused multiply-adds (single rounding each), the way nvcc
// contracts the reference's a + (b + (c + d t) t) t (cudaICEEMDAN.cu:436);
// the oracle evaluates the same fma chain.
template <typename real>
__device__ __forceinline__ real horner(const Coef<real>& q, real t)
{
    return fma(fma(fma(q.d, t, q.c), t, q.b), t, q.a);
}

// ------------------------------------------------------- windowed spline solve

// Row g of the spline system (cudaICEEMDAN.cu:310-343) from the lengths
// hl, hr of the intervals left and right of knot g and their slopes sl, sr
// ((y_g - y_{g-1}) / hl and (y_{g+1} - y_g) / hr, each computed once per
// interval and shared by the two rows it touches); rows outside [0, n) and
// the two end rows are identity rows with a zero right-hand side.
// Not-a-knot ends (NAK, n >= 4): c_0 is eliminated from row 1 and c_{n-1}
// from row n-2 (both stay diagonally dominant); the end values follow from
// nak_end after the solve.
template <typename real, bool NAK>
__device__ __forceinline__ void spline_row(int n, long g, int hl, int hr, real sl, real sr,
                                           real& dl, real& d, real& du, real& f)
{
    // computed on every lane and selected at the end (an early return here
    // kept the four outputs in scratch memory); rows outside (0, n-1) see
    // neighbours that exist in the window's LDS but take the identity row
    const bool inner = g > 0 && g < n - 1;
    real tdl = (real)hl;
    real tdu = (real)hr;
    real td = (real)(2 * (hl + hr));
    const real tf = (real)3 * (sr - sl);
    if (NAK && n >= 4) {
        if (g == 1) {
            tdl = (real)0;
            td = (real)(3 * hl + 2 * hr) + (real)hl * (real)hl / (real)hr;
            tdu = (real)hr - (real)hl * (real)hl / (real)hr;
        }
        if (g == n - 2) {
            tdl = (real)hl - (real)hr * (real)hr / (real)hl;
            td = (real)(2 * hl + 3 * hr) + (real)hr * (real)hr / (real)hl;
            tdu = (real)0;
        }
    }
    dl = inner ? tdl : (real)0;
    d = inner ? td : (real)1;
    du = inner ? tdu : (real)0;
    f = inner ? tf : (real)0;
}

// not-a-knot end value: c_e = c_1 (1 + r) - c_2 r, r = h_e / h_next
template <typename real>
__device__ __forceinline__ real nak_end(real c1, real c2, int h_e, int h_next)
{
    const real r = (real)h_e / (real)h_next;
    return c1 * ((real)1 + r) - c2 * r;
}

// Positions of knots [glo, ghi) of (series s, type) into LDS (sp index =
// g - lo), from the tile bitmasks (mask [tile][type][row]), knot prefixes and
// counts.  One wavefront; tpre (64 ints) and tmask (64 x 4 words) are LDS
// scratch.  Tiles are taken 64 at a time (their prefixes and masks in LDS).
// When the knots of a chunk lie in few tiles (dense knots), every lane tests
// its own bit of each tile row and the set bits' ranks give the knot
// indices; otherwise each lane finds the tile of its knots g + lane + 64 j by
// binary search and the knot's bit by selection.  hint = the tile holding
// knot glo, or -1 (64-ary search).  NIT >= the knots of one window / 64.
// (inlined: as a called function it cost the solve 8 % and the adjoint
// window solve 1 %)
// row masks of one staged tile in LDS (TW words per tile).  The adjoint
// pads them from 4 to 5 words: the 64 lanes store their tiles' rows at a
// 40-byte stride instead of 32, which spreads the stores over the banks.
// The solve keeps 4: 2 KB less LDS per workgroup gives its FP32 instance a
// sixth wave per SIMD, which is worth more than the conflicts (-2.5 %).
constexpr int kTMaskW = kWR + 1;

template <int TW>
__device__ __forceinline__ void gather_positions(const SeriesDev& sd, const uint64_t* kmask,
                                                 const int* kpre, const int* kcnt, int s, int type,
                                                 long lo, int glo, int ghi, int* sp, int* tpre,
                                                 uint64_t (*tmask)[TW], int hint)
{
    const int lane = threadIdx.x & 63;
    const int tb = load_scalar(sd.tile0 + s), ntl = load_scalar(sd.ntl + s);
    int a = hint >= 0 ? hint : 0, b = hint >= 0 ? hint + 1 : ntl;
    while (b - a > 1) {
        const int step = (b - a + kWave - 1) / kWave;
        const int idx = a + lane * step;
        const bool ok = idx < b && kpre[2 * (tb + idx) + type] <= glo;
        const uint64_t m = wave_ballot(ok);
        const int L = 63 - __builtin_clzll(m);
        a = a + L * step;
        b = min(a + step, b);
    }
    int g = glo;
    int tile = a;
    while (g < ghi) {
        const int tl = tile + lane;
        const bool valid = tl < ntl;
        const int p = valid ? kpre[2 * (tb + tl) + type] : 0x7FFFFFFF;
        const int c = valid ? kcnt[2 * (tb + tl) + type] : 0;
        uint64_t mk[kWR];
#pragma unroll
        for (int r = 0; r < kWR; ++r) mk[r] = valid ? kmask[8 * (int64_t)(tb + tl) + kWR * type + r] : 0ull;
        tpre[lane] = p;
#pragma unroll
        for (int r = 0; r < kWR; ++r) tmask[lane][r] = mk[r];
        const uint64_t vm = wave_ballot(valid);
        if (vm == 0ull) break;   // counts inconsistent with n: leave the rest
        const int last = 63 - __builtin_clzll(vm);
        const int pend = __shfl(p + c, last, kWave);
        const int gend = min(ghi, pend);
        // chunk tiles holding the first and the last knot of [g, gend)
        const uint64_t ua = vm & wave_ballot(p <= g), ub = vm & wave_ballot(p <= gend - 1);
        const int xa = ua ? 63 - __builtin_clzll(ua) : 0;
        const int xb = ub ? 63 - __builtin_clzll(ub) : 0;
        wave_sync();
        if (xb - xa < 16) {
            // few tiles with many knots each: the wave goes through the
            // tiles, one lane per sample of a row
            for (int x = xa; x <= xb; ++x) {
                // the tile's prefix and four row masks in one round of LDS
                // reads, moved to scalar registers (the same in every lane):
                // the lane test is then the exec mask, the counts scalar
                int cb = __builtin_amdgcn_readfirstlane(tpre[x]);
                uint64_t mr[kWR];
#pragma unroll
                for (int r = 0; r < kWR; ++r) {
                    const uint64_t m = tmask[x][r];
                    mr[r] = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(m >> 32)) << 32) |
                            (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)m);
                }
#pragma unroll
                for (int r = 0; r < kWR; ++r) {
                    if (__builtin_amdgcn_inverse_ballot_w64(mr[r])) {
                        const int gi = __builtin_amdgcn_mbcnt_hi(
                            (uint32_t)(mr[r] >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mr[r], (uint32_t)cb));
                        if (gi >= g && gi < gend) sp[gi - lo] = (tile + x) * kWT + kWave * r + lane;
                    }
                    cb += __popcll(mr[r]);
                }
            }
        } else if (lane >= xa && lane <= xb) {
            // many tiles with few knots each: every lane walks the set bits
            // of its own tile (prefix and row masks still in its registers)
            int gi = p;
#pragma unroll
            for (int r = 0; r < kWR; ++r) {
                uint64_t m = mk[r];
                while (m != 0ull && gi < gend) {
                    const int bit = __builtin_ctzll(m);
                    m &= m - 1ull;
                    if (gi >= g) sp[gi - lo] = (tile + lane) * kWT + kWave * r + bit;
                    ++gi;
                }
            }
        }
        g = gend;
        tile += kWave;
        wave_sync();
    }
}

// Values of knots [glo, ghi) (positions already in sp) from the series' pass
// input hs (N samples) into sy; every load is issued before any store.
template <typename real, int NIT>
__device__ void gather_values(const real* hs, int N, long lo, int glo, int ghi, const int* sp,
                              real* sy)
{
    const int lane = threadIdx.x & 63;
    for (int i0 = (int)(glo - lo); i0 < (int)(ghi - lo); i0 += kWave * NIT) {
        int pos[NIT];
#pragma unroll
        for (int j = 0; j < NIT; ++j) {
            const int i = i0 + lane + kWave * j;
            pos[j] = i < ghi - lo ? min(max(sp[i], 0), N - 1) : -1;
        }
        real yv[NIT];
#pragma unroll
        for (int j = 0; j < NIT; ++j) yv[j] = pos[j] >= 0 ? hs[pos[j]] : (real)0;
#pragma unroll
        for (int j = 0; j < NIT; ++j) {
            const int i = i0 + lane + kWave * j;
            if (pos[j] >= 0) sy[i] = yv[j];
        }
    }
}

// x / y in the solve: IEEE division in FP64 (the oracle's arithmetic, bit
// for bit); in FP32 one hardware reciprocal (1 ulp) and a product
template <typename real>
__device__ __forceinline__ real sdiv(real x, real y)
{
    if constexpr (std::is_same<real, double>::value) return x / y;
    else return x * (real)__builtin_amdgcn_rcpf((float)y);
}

// Shuffles of the reduction by ST lanes: lane i gets lane i - ST (lane_up)
// or i + ST (lane_down); what a lane with no source gets is unspecified (the
// callers do not use it).  By one lane a DPP whole-wave shift (one vector
// move per 32 bits); farther a ds_bpermute.
template <int CTRL>
__device__ __forceinline__ int dpp_move(int v) { return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, true); }
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) { return __int_as_float(dpp_move<CTRL>(__float_as_int(v))); }
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = dpp_move<CTRL>((int)b), hi = dpp_move<CTRL>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <int ST, typename real>
__device__ __forceinline__ real lane_up(real v)
{
    if constexpr (ST == 1) return dpp_move<0x138>(v);
    else return shfl_up(v, ST);
}
template <int ST, typename real>
__device__ __forceinline__ real lane_down(real v)
{
    if constexpr (ST == 1) return dpp_move<0x130>(v);
    else return shfl_down(v, ST);
}

template <typename real, int B, class RowFn>
__device__ __forceinline__ void partition_pcr(int lane, RowFn row, real (&x)[B])
{
    real cp[B], al[B], be[B], ga[B];
#pragma unroll
    for (int q = 0; q <= B - 2; ++q) {
        real dl, d, du, f;
        row(q, dl, d, du, f);
        real inv;
        if (q == 0) {
            const real dl0 = (lane == 0) ? (real)0 : dl;
            inv = sdiv<real>((real)1, d);
            al[0] = f * inv;
            be[0] = (-dl0) * inv;
        } else {
            inv = sdiv<real>((real)1, d - dl * cp[q - 1]);
            al[q] = (f - dl * al[q - 1]) * inv;
            be[q] = (-dl) * be[q - 1] * inv;
        }
        cp[q] = du * inv;
        ga[q] = (q == B - 2) ? (-du) * inv : (real)0;
    }
#pragma unroll
    for (int q = B - 3; q >= 0; --q) {
        al[q] = al[q] - cp[q] * al[q + 1];
        be[q] = be[q] - cp[q] * be[q + 1];
        ga[q] = ga[q] - cp[q] * ga[q + 1];
    }
    // interface row q = B-1 -> reduced tridiagonal system in z_lane
    real dl, d, du, f;
    row(B - 1, dl, d, du, f);
    real na0 = lane_down<1>(al[0]), nb0 = lane_down<1>(be[0]), ng0 = lane_down<1>(ga[0]);
    if (lane == kWave - 1) {
        du = (real)0;
        na0 = (real)0;
        nb0 = (real)0;
        ng0 = (real)0;
    }
    real A = dl * be[B - 2];
    real D = (d + dl * ga[B - 2]) + du * nb0;
    real C = du * ng0;
    real R = (f - dl * al[B - 2]) - du * na0;
    auto level = [&](auto st_c) {
        constexpr int st = decltype(st_c)::value;
        const real Am = lane_up<st>(A), Dm = lane_up<st>(D), Cm = lane_up<st>(C),
                   Rm = lane_up<st>(R);
        const real Ap = lane_down<st>(A), Dp = lane_down<st>(D), Cp = lane_down<st>(C),
                   Rp = lane_down<st>(R);
        real a2 = (real)0, c2 = (real)0, d2 = D, r2 = R;
        if (lane - st >= 0) {
            const real k1 = sdiv<real>(A, Dm);
            a2 = -(Am * k1);
            d2 = d2 - Cm * k1;
            r2 = r2 - Rm * k1;
        }
        if (lane + st < kWave) {
            const real k2 = sdiv<real>(C, Dp);
            c2 = -(Cp * k2);
            d2 = d2 - Ap * k2;
            r2 = r2 - Rp * k2;
        }
        A = a2;
        C = c2;
        D = d2;
        R = r2;
    };
    level(std::integral_constant<int, 1>{});
    level(std::integral_constant<int, 2>{});
    level(std::integral_constant<int, 4>{});
    level(std::integral_constant<int, 8>{});
    level(std::integral_constant<int, 16>{});
    level(std::integral_constant<int, 32>{});
    const real z = sdiv<real>(R, D);
    real zl = lane_up<1>(z);
    if (lane == 0) zl = (real)0;
#pragma unroll
    for (int q = 0; q < B; ++q) x[q] = (q == B - 1) ? z : (al[q] + be[q] * zl) + ga[q] * z;
}

// One wavefront per window item.  NAK = not-a-knot ends, as its own
// instance: the natural-end instance carries no end-row divisions (FP64:
// 113 registers, four waves per SIMD; the not-a-knot one 149, three).
template <typename real, int B, int H, bool NAK>
__global__ __launch_bounds__(kThreads) void k_solve(SolveArgs<real> a)
{
    constexpr int W = kWave * B;
    constexpr int K = W - 2 * H;
    constexpr int NW = kThreads / kWave;
    // rows padded to a multiple of 16 bytes, so a lane's B + 2 entries at
    // r = lane * B start 16-byte aligned
    __shared__ __align__(16) int s_pos[NW][W + 4];
    __shared__ __align__(16) real s_y[NW][W + 4];
    __shared__ int s_tpre[NW][kWave];
    __shared__ uint64_t s_tmask[NW][kWave][kWR];
    const int wv = wave_index(), lane = threadIdx.x & 63;
    int* sp = s_pos[wv];
    real* sy = s_y[wv];
    // window items and series metadata through the scalar cache (written
    // by earlier launches): one round trip for the item, one for its series
    if (a.run && load_scalar(a.run) == 0u) return;
    const unsigned total = load_scalar(a.wincount);
    for (unsigned item = blockIdx.x * NW + wv; item < total; item += gridDim.x * NW) {
        const int s = load_scalar(a.winlist + kWinItem * item);
        const int code = load_scalar(a.winlist + kWinItem * item + 1);
        const int hint = load_scalar(a.winlist + kWinItem * item + 2);
        const int type = code & 1;
        const int w = code >> 1;
        if (a.active && !load_scalar(a.active + s)) continue;
        const int n = load_scalar(a.k.n + 2 * s + type);
        const int64_t kb = load_scalar(a.sd.kbase + s) + (type ? a.sd.kstride : 0);
        const long g0 = (long)w * K - H;
        const long lo = g0 - 1;
        const int glo = (int)max(lo, 0L);
        const int ghi = (int)min(lo + W + 2, (long)n);
        constexpr int NIT = (W + 2 + kWave - 1) / kWave;
        gather_positions<kWR>(a.sd, a.k.mask, a.k.pre, a.k.cnt, s, type, lo, glo, ghi, sp,
                              s_tpre[wv], s_tmask[wv], hint);
        wave_sync();
        gather_values<real, NIT>(a.h + load_scalar(a.sd.off + s), load_scalar(a.sd.len + s), lo, glo,
                                 ghi, sp, sy);
        wave_sync();
        // endpoint fix-up when n >= 6 (setBoundaryMax/Min, cudaICEEMDAN.cu:209-291);
        // idempotent, so every window that holds an end applies it
        if (lane == 0) {
            const long keep0 = (long)w * K;
            if (lo <= 0) {
                const int i0 = (int)(-lo);
                int replaced = 0;
                if (n >= 6) {
                    const real slope =
                        (sy[i0 + 2] - sy[i0 + 1]) / (real)(sp[i0 + 2] - sp[i0 + 1]);
                    const real t = sy[i0 + 1] - slope * (real)(sp[i0 + 1] - sp[i0]);
                    if (type == 0 ? (t > sy[i0]) : (t < sy[i0])) {
                        sy[i0] = t;
                        replaced = 1;
                    }
                }
                if (w == 0) {
                    a.k.fix[4 * s + 2 * type] = sy[i0];
                    if (a.fixf) a.fixf[4 * s + 2 * type] = replaced;
                }
            }
            const long ie = (long)(n - 1) - lo;
            if (ie >= 0 && ie < W + 2) {
                const int i1 = (int)ie;
                int replaced = 0;
                if (n >= 6) {
                    const real slope =
                        (sy[i1 - 1] - sy[i1 - 2]) / (real)(sp[i1 - 1] - sp[i1 - 2]);
                    const real t = sy[i1 - 1] + slope * (real)(sp[i1] - sp[i1 - 1]);
                    if (type == 0 ? (t > sy[i1]) : (t < sy[i1])) {
                        sy[i1] = t;
                        replaced = 1;
                    }
                }
                if (n - 1 >= keep0 && n - 1 < keep0 + K) {
                    a.k.fix[4 * s + 2 * type + 1] = sy[i1];
                    if (a.fixf) a.fixf[4 * s + 2 * type + 1] = replaced;
                }
            }
        }
        wave_sync();

        // the lane's rows r = rb .. rb+B-1 (LDS index r + 1) need window
        // knots rb .. rb+B+1: loaded once, as wide reads, and the B + 1
        // intervals between them give every row its lengths and slopes
        const int rb = lane * B;
        int hx[B + 1];
        real sl[B + 1];
        {
            int px[B + 2];
            real py[B + 2];
#pragma unroll
            for (int j = 0; j < B + 2; ++j) {
                px[j] = sp[rb + j];
                py[j] = sy[rb + j];
            }
#pragma unroll
            for (int j = 0; j <= B; ++j) {
                hx[j] = px[j + 1] - px[j];
                sl[j] = sdiv<real>(py[j + 1] - py[j], (real)hx[j]);
            }
        }
        real xs[B];
        partition_pcr<real, B>(
            lane,
            [&](int q, real& dl, real& d, real& du, real& f) {
                spline_row<real, NAK>(n, g0 + rb + q, hx[q], hx[q + 1], sl[q], sl[q + 1], dl, d, du,
                                      f);
            },
            xs);
        real* cout = a.c + kb;
        const bool nak = NAK && n >= 4;
#pragma unroll
        for (int q = 0; q < B; ++q) {
            const int r = rb + q;
            const long gg = g0 + r;
            if (r >= H && r < H + K && gg < n) {
                if (!(nak && (gg == 0 || gg == n - 1))) cout[gg] = xs[q];
            }
        }
        if (nak) {
            // the window that keeps an end knot recovers it from its own rows
            wave_sync();
#pragma unroll
            for (int q = 0; q < B; ++q) sy[rb + q + 1] = xs[q];
            wave_sync();
            if (lane == 0) {
                const long r0 = -g0, r1 = (long)(n - 1) - g0;
                if (r0 >= H && r0 < H + K) {
                    const int i0 = (int)r0 + 1;
                    cout[0] = nak_end<real>(sy[i0 + 1], sy[i0 + 2], sp[i0 + 1] - sp[i0],
                                            sp[i0 + 2] - sp[i0 + 1]);
                }
                if (r1 >= H && r1 < H + K) {
                    const int i1 = (int)r1 + 1;
                    cout[n - 1] = nak_end<real>(sy[i1 - 1], sy[i1 - 2], sp[i1] - sp[i1 - 1],
                                                sp[i1 - 1] - sp[i1 - 2]);
                }
            }
        }
        wave_sync();
    }
}

// ------------------------------------------------------------ Rilling decide

// siftingCriterion (cudaICEEMDAN.cu:494-525) + checkBreak (:546-580) with
// per-series masks: export at the first stop or at pass cap-1.
template <typename real>
__global__ void k_rilling_decide(SeriesDev sd, const int* nk, const int* sx_cnt,
                                 const int* sx_mark, int* active, int* stop_pass, int* subs,
                                 int pass, int cap, const unsigned* run, unsigned* n_active)
{
    const int s = blockIdx.x;
    const int lane = threadIdx.x;
    if (load_scalar(run) == 0u) return;
    if (!active[s]) return;
    long long c = 0;
    int mk = 0;
    const int t0 = sd.tile0[s], nt = sd.ntl[s];
    for (int t = lane; t < nt; t += kWave) {
        c += sx_cnt[t0 + t];
        mk |= sx_mark[t0 + t];
    }
    for (int o = 32; o >= 1; o >>= 1) {
        c += __shfl_xor(c, o, kWave);
        mk |= __shfl_xor(mk, o, kWave);
    }
    if (lane == 0) {
        const int N = sd.len[s];
        const real frac = (real)c * (real)(1.0 / (double)N);
        const bool cont = (frac > (real)0.05) || (mk && (nk[2 * s] + nk[2 * s + 1]) > 6);
        if (!cont || pass == cap - 1) {
            active[s] = 0;
            stop_pass[s] = pass;
        } else {
            subs[s] += 1;
            atomicAdd(n_active, 1u);
        }
    }
}

// --------------------------------------------------------------- statistics

// Canonical sum: lane l of 1024 adds x[l], x[l+1024], ... in order, then a
// pairwise tree (same order as the oracle).  One workgroup per series.
template <typename real>
__device__ real canon_sum_block(real acc, real* sred)
{
    const int tid = threadIdx.x;
    sred[tid] = acc;
    __syncthreads();
    for (int st = 512; st >= 1; st >>= 1) {
        if (tid < st) sred[tid] = sred[tid] + sred[tid + st];
        __syncthreads();
    }
    const real r = sred[0];
    __syncthreads();
    return r;
}

template <typename real>
__global__ __launch_bounds__(1024) void k_series_var(const real* base, const int64_t* off,
                                                     int64_t stride, const int* len, real* var_out)
{
    __shared__ real sred[1024];
    const int s = blockIdx.x;
    const int64_t o = off ? off[s] : (int64_t)s * stride;
    const real* x = base + o;
    const int N = len ? len[s] : (int)stride;
    real acc = (real)0;
    for (int p = threadIdx.x; p < N; p += 1024) acc = acc + x[p];
    const real mu = canon_sum_block<real>(acc, sred) * (real)(1.0 / (double)N);
    acc = (real)0;
    for (int p = threadIdx.x; p < N; p += 1024) {
        const real dv = x[p] - mu;
        acc = acc + dv * dv;
    }
    const real ss = canon_sum_block<real>(acc, sred);
    if (threadIdx.x == 0) var_out[s] = ss * (real)(1.0 / (double)(N - 1));
}

// ------------------------------------------------- sift pass, wave per tile
// One wavefront sifts a run of consecutive 256-sample tiles of one series;
// lane l owns samples t0 + 64 r + l, r = 0..3.  Everything a tile needs is
// wave-local: the knot flags of h are ballots (their ranks give each
// sample's interval), the knot table [pre-2, pre+cnt] of both types lives in
// the wave's own LDS slice, and no workgroup barrier is used, so the four
// waves of a workgroup run independent segments.

// base + the number of set bits of `mask` below this lane (two mbcnt steps;
// running bases instead of a per-row prefix array keep eight scalar
// registers free, which the sift loop would otherwise spill)
__device__ __forceinline__ int rank_below(uint64_t mask, int base)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, (uint32_t)base));
}

// Lane 0 stores the eight wave-uniform row masks m[type][row] (no lane-
// dependent indexing, which would put the array in scratch memory).
__device__ __forceinline__ void store_masks(uint64_t* dst, const uint64_t (&m)[2][kWR], int lane)
{
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < 2 * kWR; ++q) dst[q] = m[q / kWR][q % kWR];
    }
}

// Per-wave knot table of a tile (KC entries per type): each knot's position
// and the coefficients of the interval that starts at it.  The small
// tables keep one record per knot (coefficients, then the position: 32
// bytes in FP32, 48 in FP64), so that a knot's position and coefficients
// are read at one LDS address (-2.5 % on the FP32 sift, -1 % on FP64); the
// full tables keep two arrays (records would cost them waves per SIMD:
// their LDS would grow from 22 to 35 KB per workgroup in FP32, to 52 KB in
// FP64).
template <typename real, int KC, bool REC>
struct KnotTableOf {
    int pos[2][KC];
    Coef<real> q[2][KC];
    __device__ __forceinline__ int& P(int ty, int j) { return pos[ty][j]; }
    __device__ __forceinline__ Coef<real>& Q(int ty, int j) { return q[ty][j]; }
};
template <typename real, int KC>
struct KnotTableOf<real, KC, true> {
    struct alignas(sizeof(real) == 4 ? 32 : 16) Rec {
        Coef<real> q;
        int pos;
    };
    Rec r[2][KC];
    __device__ __forceinline__ int& P(int ty, int j) { return r[ty][j].pos; }
    __device__ __forceinline__ Coef<real>& Q(int ty, int j) { return r[ty][j].q; }
};
template <typename real, int KC>
using KnotTable = KnotTableOf<real, KC, KC == kWKS>;

template <typename real, int KC>
struct WTile {
    static constexpr int NC = (KC + kWave - 1) / kWave;
    real h[kWR];
    real hl, hr;          // h[t0 - 1], h[t0 + kWT] (0 outside the series)
    real c[2][NC];        // c of knots pre-2+j, j = lane + 64 k
    int nbx;              // lanes 0..5: positions of knots pre-1, pre-2, pre+cnt per type
    real nby;
    uint64_t mk;          // lanes 0..7: knot bitmasks [type][row] of h (other lanes: copies)
};


// One segment of tiles, sifted by one wavefront with a knot table of KC
// entries per type.  A tile whose knots do not fit (cnt + 3 > KC) is an error
// with the full table (KC = kWKC fits every tile); with the small table it is
// appended to a.ovf_segs for a second launch with the full table, and the
// tile after it takes h_new of its left neighbour from its own knots.
template <typename real, bool RILL, int KC>
__device__ __forceinline__ void sift_segment(const SiftArgs<real>& a, const Segment* segs, int sid,
                                             int lane, KnotTable<real, KC>& T)
{
    const SeriesDev& sd = a.sd;
    const Segment seg = segs[sid];
    const int s = seg.s;
    if (a.active && !a.active[s]) return;
    const int N = sd.len[s];
    const int tb = sd.tile0[s];
    const int64_t kb0 = sd.kbase[s], kb1 = sd.kstride + sd.kbase[s];
    const real* h = a.h_in + sd.off[s];
    real* hout = a.h_out + sd.off[s];
    // FP32: the two sample pointers in vector registers (the empty asm makes
    // them non-uniform), which frees four scalar registers through the loop;
    // the small-table instance then spills 8 instead of 21 (FP64 has no
    // vector registers to spare)
    if (std::is_same<real, float>::value) asm volatile("" : "+v"(h), "+v"(hout));
    const int nn[2] = {a.kin.n[2 * s], a.kin.n[2 * s + 1]};
    const real fx[2][2] = {{a.kin.fix[4 * s], a.kin.fix[4 * s + 1]},
                           {a.kin.fix[4 * s + 2], a.kin.fix[4 * s + 3]}};
    const bool use_mean = (nn[0] > 3) && (nn[1] > 3);

    auto load = [&](int t, const int* pre, const int* cnt, WTile<real, KC>& w) {
        const int t0 = t * kWT;
        if (std::is_same<real, float>::value && t0 + kWT <= N) {
            // a whole tile (every one but a series' last): no lane tests
            // (FP32 only: the FP64 instances came out 0.8 % slower with it)
#pragma unroll
            for (int r = 0; r < kWR; ++r) w.h[r] = h[t0 + r * kWave + lane];
        } else {
#pragma unroll
            for (int r = 0; r < kWR; ++r) {
                const int p = t0 + r * kWave + lane;
                w.h[r] = p < N ? h[p] : (real)0;
            }
        }
        w.hl = t0 >= 1 ? h[t0 - 1] : (real)0;
        w.hr = t0 + kWT < N ? h[t0 + kWT] : (real)0;
#pragma unroll
        for (int ty = 0; ty < 2; ++ty) {
            const real* cp = a.c + (ty ? kb1 : kb0);
#pragma unroll
            for (int k = 0; k < WTile<real, KC>::NC; ++k) {
                const int j = lane + k * kWave;
                const int g = pre[ty] - 2 + j;
                w.c[ty][k] = (j < cnt[ty] + 3 && g >= 0 && g < nn[ty]) ? cp[g] : (real)0;
            }
        }
        w.nbx = -1;
        w.nby = (real)0;
        if (lane < 6) {
            const int64_t o = 6 * (int64_t)(tb + t) + lane;
            w.nbx = a.kin.nbx[o];
            w.nby = a.kin.nby[o];
        }
        w.mk = a.kin.mask[8 * (int64_t)(tb + t) + (lane & (2 * kWR - 1))];
    };

    int mpre[2], mcnt[2], npre[2] = {0, 0}, ncnt[2] = {0, 0};
    {
        const int gt = tb + seg.t_first;
        mpre[0] = load_scalar(a.kin.pre + 2 * gt);
        mpre[1] = load_scalar(a.kin.pre + 2 * gt + 1);
        mcnt[0] = load_scalar(a.kin.cnt + 2 * gt);
        mcnt[1] = load_scalar(a.kin.cnt + 2 * gt + 1);
    }
    WTile<real, KC> cur, nxt;
    load(seg.t_first, mpre, mcnt, cur);
    // the first tile is needed at once; waiting here also tells the compiler
    // that nothing loaded before the loop is outstanding inside it, so the
    // loop body never waits on the prefetch of the next tile
    wait_vmem();
    // h_new of the sample just before the tile: the previous tile of the
    // segment computed it (same interval, same coefficients, same bits)
    real hcarry = (real)0;
    bool carry = false;   // hcarry is valid (the previous tile was sifted here)

    for (int i = 0; i < seg.count; ++i) {
        const int t = seg.t_first + i;
        const int gt = tb + t;
        const int t0 = t * kWT;
        const int nt = min(kWT, N - t0);
        const int pre[2] = {mpre[0], mpre[1]}, cnt[2] = {mcnt[0], mcnt[1]};
        if (i + 1 < seg.count) {
            // the next tile's prefixes and counts, then its prefetch (which
            // needs them for its coefficient loads); loading them a tile
            // earlier hid the scalar loads' latency but held four more scalar
            // registers through the loop, which spilled (through vector
            // lanes) at 5 % of the pass's time
            const int g1 = gt + 1;
            npre[0] = load_scalar(a.kin.pre + 2 * g1);
            npre[1] = load_scalar(a.kin.pre + 2 * g1 + 1);
            ncnt[0] = load_scalar(a.kin.cnt + 2 * g1);
            ncnt[1] = load_scalar(a.kin.cnt + 2 * g1 + 1);
            load(t + 1, npre, ncnt, nxt);
        }

        // ---- knot flags of h (find_extrema_shfl_*, cudaICEEMDAN.cu:92-148):
        // the previous pass (or the stage start) found them when it wrote h,
        // from the same neighbour values, and stored them as bitmasks
        uint64_t bm[2][kWR];
        int tot[2] = {0, 0};
#pragma unroll
        for (int ty = 0; ty < 2; ++ty)
#pragma unroll
            for (int r = 0; r < kWR; ++r) {
                bm[ty][r] = lane_value(cur.mk, ty * kWR + r);
                tot[ty] += __popcll(bm[ty][r]);
            }
        const bool over = cnt[0] + 3 > KC || cnt[1] + 3 > KC;
        if (tot[0] != cnt[0] || tot[1] != cnt[1] || (over && !a.ovf_segs)) {
            if (lane == 0) atomicOr(a.pc.err, ICM_DEVERR_KNOTCAP);
            return;
        }

        // interior tiles (neither series end inside, all 256 samples valid)
        // take a copy of the body without the per-lane end tests
        auto tile = [&](auto full_c) {
            constexpr bool FULL = decltype(full_c)::value;
            // ---- knot table: index j <-> knot pre - 2 + j
            int rb[2] = {2, 2};   // table index of the row's first knot
#pragma unroll
            for (int r = 0; r < kWR; ++r) {
                const int p = t0 + r * kWave + lane;
#pragma unroll
                for (int ty = 0; ty < 2; ++ty) {
                    if (__builtin_amdgcn_inverse_ballot_w64(bm[ty][r])) {
                        const int j = rank_below(bm[ty][r], rb[ty]);
                        const int g = pre[ty] - 2 + j;
                        real y = cur.h[r];
                        if (!FULL) {   // knots 0 and n-1 lie in the series' end tiles
                            if (g == 0) y = fx[ty][0];
                            if (g == nn[ty] - 1) y = fx[ty][1];
                        }
                        T.P(ty, j) = p;
                        T.Q(ty, j).a = y;
                    }
                    rb[ty] += __popcll(bm[ty][r]);
                }
            }
            if (lane < 6) {
                const int ty = lane / 3, w = lane % 3;
                const int j = w == 0 ? 1 : (w == 1 ? 0 : cnt[ty] + 2);
                const int g = pre[ty] - 2 + j;
                if (cur.nbx >= 0 && g >= 0 && g < nn[ty]) {
                    real y = cur.nby;
                    if (g == 0) y = fx[ty][0];
                    if (g == nn[ty] - 1) y = fx[ty][1];
                    T.P(ty, j) = cur.nbx;
                    T.Q(ty, j).a = y;
                }
            }
#pragma unroll
            for (int ty = 0; ty < 2; ++ty)
#pragma unroll
                for (int k = 0; k < WTile<real, KC>::NC; ++k) {
                    const int j = lane + k * kWave;
                    if (j < cnt[ty] + 3) T.Q(ty, j).c = cur.c[ty][k];
                }
            wave_sync();
            // ---- interval coefficients (spline_coefficients, cudaICEEMDAN.cu:394-398)
            // both types in one strided loop (intervals j of type 0, then of type 1)
            for (int q = lane; q < cnt[0] + cnt[1] + 4; q += kWave) {
                const int ty = q >= cnt[0] + 2;
                const int j = ty ? q - (cnt[0] + 2) : q;
                const int g = pre[ty] - 2 + j;
                if (g < 0 || g + 1 > nn[ty] - 1) continue;
                const real hk = (real)(T.P(ty, j + 1) - T.P(ty, j));
                const real a0 = T.Q(ty, j).a, a1 = T.Q(ty, j + 1).a;
                const real c0 = T.Q(ty, j).c, c1 = T.Q(ty, j + 1).c;
                if (std::is_same<real, double>::value) {
                    // the reference's expressions, correctly rounded divisions:
                    // FP64 equals the oracle bit for bit
                    T.Q(ty, j).b = (a1 - a0) / hk - hk * ((real)2 * c0 + c1) / (real)3;
                    T.Q(ty, j).d = (c1 - c0) / ((real)3 * hk);
                } else {
                    // FP32 (compared with the FP64 oracle by its error, not
                    // bit for bit): the hardware reciprocal of the spacing
                    // (1 ulp) and 1/3 as a product, instead of three
                    // correctly rounded divisions (~9 instructions each) --
                    // 9 % of the sift pass's time at the 64-channel FP32 job
                    const real rh = (real)__builtin_amdgcn_rcpf((float)hk);
                    const real third = (real)(1.0 / 3.0);
                    T.Q(ty, j).b = (a1 - a0) * rh - hk * ((real)2 * c0 + c1) * third;
                    T.Q(ty, j).d = (c1 - c0) * third * rh;
                }
            }
            wave_sync();

            // ---- envelopes, mean, h_new = h - m (interpolate :403-439,
            // averageUppperLower :441-454, updateRealizations :527-544)
            real hn[kWR];
            int sxc = 0;
            bool sxm = false;
            // MEAN 1: the envelopes' mean is subtracted (both have > 3
            // knots); 0: h' = h (only the Rilling statistics need U, L);
            // 2: chosen per sample by use_mean.  FP32 takes 0 or 1 by a
            // branch per tile (-1.6 % on the FP32 sift); the FP64 instances
            // came out slower with the two copies and keep 2.
            auto rows = [&](auto mean_c) {
                constexpr int MEAN = decltype(mean_c)::value;
                int eb[2] = {1, 1};   // interval index of the row's first sample, less its rank
#pragma unroll
                for (int r = 0; r < kWR; ++r) {
                    const int u = r * kWave + lane;
                    const int p = t0 + u;
                    hn[r] = cur.h[r];
                    if (MEAN == 0 && !RILL) {
                        if (FULL || u < nt) hout[p] = hn[r];
                        continue;
                    }
                    const int ru = rank_below(bm[0][r], eb[0]), rl = rank_below(bm[1][r], eb[1]);
                    eb[0] += __popcll(bm[0][r]);
                    eb[1] += __popcll(bm[1][r]);
                    if (FULL || u < nt) {
                        const bool p0 = !FULL && p == 0;
                        const int ju = p0 ? 2 : ru;
                        const int jl = p0 ? 2 : rl;
                        const real U = horner(T.Q(0, ju), (real)(p - T.P(0, ju)));
                        const real L = horner(T.Q(1, jl), (real)(p - T.P(1, jl)));
                        if (MEAN == 1) hn[r] = cur.h[r] - (U + L) / (real)2;
                        if (MEAN == 2) hn[r] = cur.h[r] - (use_mean ? (U + L) / (real)2 : (real)0);
                        hout[p] = hn[r];
                        if (RILL) {
                            const real sx = fabs((U + L) / (U - L));
                            sxc += (sx > (real)0.05) ? 1 : 0;
                            sxm |= sx > (real)0.5;
                        }
                    }
                }
            };
            if (!std::is_same<real, float>::value) rows(std::integral_constant<int, 2>{});
            else if (use_mean) rows(std::integral_constant<int, 1>{});
            else rows(std::integral_constant<int, 0>{});
            if (RILL) {
                for (int o = 32; o >= 1; o >>= 1) sxc += __shfl_xor(sxc, o, kWave);
                const uint64_t mk = wave_ballot(sxm);
                if (lane == 0) {
                    a.sx_cnt[gt] = sxc;
                    a.sx_mark[gt] = mk != 0ull ? 1 : 0;
                }
            }

            if (a.emit) {
                // h_new just outside the tile, evaluated by every lane (same LDS words)
                real hnl = hcarry, hnr = (real)0;
                if (!carry && (FULL || t0 >= 1)) {
                    const int p = t0 - 1;
                    const int ju = T.P(0, 1) == p ? 0 : 1;
                    const int jl = T.P(1, 1) == p ? 0 : 1;
                    const real U = horner(T.Q(0, ju), (real)(p - T.P(0, ju)));
                    const real L = horner(T.Q(1, jl), (real)(p - T.P(1, jl)));
                    hnl = cur.hl - (use_mean ? (U + L) / (real)2 : (real)0);
                }
                if (FULL || t0 + kWT < N) {
                    const int p = t0 + kWT;
                    const int ju = cnt[0] + 1, jl = cnt[1] + 1;
                    const real U = horner(T.Q(0, ju), (real)(p - T.P(0, ju)));
                    const real L = horner(T.Q(1, jl), (real)(p - T.P(1, jl)));
                    hnr = cur.hr - (use_mean ? (U + L) / (real)2 : (real)0);
                }
                // ---- knots of h_new (select_extrema_*, :150-207): their row
                // bitmasks and counts (the next pass's knots)
                int run[2] = {0, 0};
                uint64_t em[2][kWR];
#pragma unroll
                for (int r = 0; r < kWR; ++r) {
                    const int u = r * kWave + lane;
                    const int p = t0 + u;
                    const real lft = r == 0 ? shift_prev(hn[0], hnl) : row_prev(hn[r], hn[r > 0 ? r - 1 : 0]);
                    const real rgt = r == kWR - 1 ? shift_next(hn[r], hnr)
                                                  : row_next(hn[r], hn[r < kWR - 1 ? r + 1 : r]);
                    row_extrema<FULL>(hn[r], lft, rgt, FULL ? 0ull : wave_ballot(u < nt),
                                      FULL ? 0ull : wave_ballot(p == 0) | wave_ballot(p == N - 1),
                                      em[0][r], em[1][r]);
#pragma unroll
                    for (int ty = 0; ty < 2; ++ty) run[ty] += __popcll(em[ty][r]);
                }
                store_masks(a.kout.mask + 8 * (int64_t)gt, em, lane);
                if (lane == 0) {
                    a.kout.cnt[2 * gt] = run[0];
                    a.kout.cnt[2 * gt + 1] = run[1];
                }
            }
            hcarry = lane_value(hn[kWR - 1], kWave - 1);
            carry = true;
        };
        if (over) {
            // left to the launch with the full table
            if (lane == 0) {
                const unsigned q = atomicAdd(a.ovf_count, 1u);
                Segment o;
                o.s = s;
                o.t_first = t;
                o.count = 1;
                o.pad = 0;
                a.ovf_segs[q] = o;
            }
            carry = false;
        } else if (t0 > 0 && t0 + kWT < N) {
            tile(std::true_type{});
        } else {
            tile(std::false_type{});
        }
        // the table is rewritten by the next tile
        wave_sync();
        cur = nxt;
        mpre[0] = npre[0];
        mpre[1] = npre[1];
        mcnt[0] = ncnt[0];
        mcnt[1] = ncnt[1];
    }
}

// One wavefront per segment (QUEUE = false), or (QUEUE = true) a resident
// grid of wavefronts working through the pass's segments: each wave takes
// the next segment from the pass's queue (one atomic increment, requested
// while the previous segment is sifted, so its latency overlaps that
// segment's loads) until the queue is exhausted.  The queue instance keeps
// the work of a launch independent of its grid (any number of segments);
// its loop keeps ~15 more registers live than the one-segment body, so it is
// used only where the one-wave-per-segment grid would exceed a launch's
// 2^32 work-items (or when ICM_SIFT_QUEUE=1).  Workgroups of the
// one-segment grid are dispatched as slots free up, which balances uneven
// segments as well.  The small-table instances are held to 96 registers, so
// that five of their waves fit on a SIMD (their tables allow eight).
template <typename real, bool RILL, int KC, bool QUEUE>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(KC == kWKS ? 5 : 1)))
void k_sift_w(SiftArgs<real> a, int nsegs)
{
    constexpr int NW = kThreads / kWave;
    __shared__ KnotTable<real, KC> s_tab[NW];
    const int lane = threadIdx.x & 63, wv = wave_index();
    if (a.pc.run && load_scalar(a.pc.run) == 0u) return;
    // a small-table pass whose input has too many over-full tiles is left
    // to k_sift_list
    if (KC == kWKS && load_scalar(a.dense) > a.dense_max) return;
    if (!QUEUE) {
        const int sid = blockIdx.x * NW + wv;
        if (sid < nsegs) sift_segment<real, RILL, KC>(a, a.segs, sid, lane, s_tab[wv]);
        return;
    }
    unsigned got = lane == 0 ? atomicAdd(a.queue, 1u) : 0u;
    int sid = __builtin_amdgcn_readlane((int)got, 0);
    while (sid < nsegs) {
        const unsigned next = lane == 0 ? atomicAdd(a.queue, 1u) : 0u;
        sift_segment<real, RILL, KC>(a, a.segs, sid, lane, s_tab[wv]);
        wave_sync();
        sid = __builtin_amdgcn_readlane((int)next, 0);
    }
}

// Second launch of a small-table pass, with full tables, as a grid-stride
// loop: the tiles the small-table launch listed (*a.ovf_count one-tile
// segments in a.ovf_segs) or, when the pass input had too many over-full
// tiles for it to run, every segment of the pass.
template <typename real, bool RILL>
__global__ __launch_bounds__(kThreads) void k_sift_list(SiftArgs<real> a, int nsegs)
{
    constexpr int NW = kThreads / kWave;
    __shared__ KnotTable<real, kWKC> s_tab[NW];
    const int lane = threadIdx.x & 63, wv = wave_index();
    if (a.pc.run && load_scalar(a.pc.run) == 0u) return;
    const bool all = load_scalar(a.dense) > a.dense_max;
    const Segment* segs = all ? a.segs : a.ovf_segs;
    const int ns = all ? nsegs : (int)load_scalar(a.ovf_count);
    SiftArgs<real> b = a;
    b.ovf_segs = nullptr;   // full tables hold every tile
    for (int sid = blockIdx.x * NW + wv; sid < ns; sid += gridDim.x * NW)
        sift_segment<real, RILL, kWKC>(b, segs, sid, lane, s_tab[wv]);
}

// Stage start for 256-sample tiles: forms x_i (the same arithmetic as
// init_value) and stores its knots as bitmasks.  MODE is ia.mode as a
// compile-time constant: one straight-line body per mode.
//
// Per-series pointers and scales of the tile being formed.
template <typename real, int MODE>
struct InitSeries {
    using P = typename Promote<real>::type;
    int tile0, tend, N;     // first tile, one past the last, samples
    const real* src;        // INIT_COPY / INIT_KNOTS / INIT_RESID
    const real* base;       // INIT_STAGE0 / INIT_STAGEK: the channel's signal
    const real* E;
    P scl0;
    real sclk;
    real* out;
    real* resid;
};

template <typename real, int MODE>
__device__ __forceinline__ InitSeries<real, MODE> init_series(const SeriesDev& sd,
                                                             const InitArgs<real>& ia,
                                                             real* h_out, int s)
{
    using P = typename Promote<real>::type;
    InitSeries<real, MODE> q{};
    q.tile0 = sd.tile0[s];
    q.tend = q.tile0 + sd.ntl[s];
    q.N = sd.len[s];
    const int64_t off = sd.off[s];
    q.out = (MODE == INIT_KNOTS || !h_out) ? nullptr : h_out + off;
    q.resid = MODE == INIT_RESID ? ia.resid + off : nullptr;
    if (MODE == INIT_COPY || MODE == INIT_KNOTS) {
        q.src = ia.src + off;
    } else if (MODE == INIT_RESID) {
        q.src = ia.src + off;
        q.E = ia.E + off;
    } else {
        const int c = sd.chan[s];
        q.base = ia.chan_sig + ia.chan_off[c];
        q.E = ia.E + sd.eoff[s];
        if (MODE == INIT_STAGE0) q.scl0 = (P)ia.scale[sd.nidx[s]];
        else q.sclk = (real)ia.scale[c];
    }
    return q;
}

// x_i of one tile (lane l: samples t0 + 64 r + l) and the samples just
// outside it; loads at clamped indices, then selects (a load under a lane
// condition would sit in its own block, waited for at its end)
template <typename real>
struct InitTile {
    real x[kWR];
    real xl, xr;
};

template <typename real, int MODE>
__device__ __forceinline__ void init_load(const InitSeries<real, MODE>& q, const InitArgs<real>& ia,
                                          int gt, int lane, InitTile<real>& w)
{
    using P = typename Promote<real>::type;
    const int t0 = (gt - q.tile0) * kWT;
    const int nt = min(kWT, q.N - t0);
    auto val = [&](int p) -> real {
        if (MODE == INIT_COPY || MODE == INIT_KNOTS) return q.src[p];
        if (MODE == INIT_RESID) return q.src[p] - q.E[p];
        const real e = q.E[p];
        if (MODE == INIT_STAGE0) return (real)((P)q.base[p] + (P)(ia.eps0 * e) * q.scl0);
        return q.base[p] + (ia.eps0 * e) * q.sclk;
    };
#pragma unroll
    for (int r = 0; r < kWR; ++r) {
        const int u = r * kWave + lane;
        const real v = val(min(t0 + u, q.N - 1));
        w.x[r] = u < nt ? v : (real)0;
    }
    const real xlv = val(max(t0 - 1, 0)), xrv = val(min(t0 + kWT, q.N - 1));
    w.xl = t0 >= 1 ? xlv : (real)0;
    w.xr = t0 + kWT < q.N ? xrv : (real)0;
}

// stores of one tile: x_i (and the noise residual), its knot bitmasks and counts
template <typename real, int MODE>
__device__ __forceinline__ void init_store(const InitSeries<real, MODE>& q, const KnotsDev<real>& kout,
                                           int gt, int lane, const InitTile<real>& w)
{
    const int t0 = (gt - q.tile0) * kWT;
    const int nt = min(kWT, q.N - t0);
#pragma unroll
    for (int r = 0; r < kWR; ++r) {
        const int u = r * kWave + lane;
        if (u < nt) {
            if (MODE == INIT_RESID ? q.out != nullptr : MODE != INIT_KNOTS) q.out[t0 + u] = w.x[r];
            if (MODE == INIT_RESID) q.resid[t0 + u] = w.x[r];
        }
    }
    int run[2] = {0, 0};
    uint64_t em[2][kWR];
#pragma unroll
    for (int r = 0; r < kWR; ++r) {
        const int u = r * kWave + lane;
        const int p = t0 + u;
        const real lft = r == 0 ? shift_prev(w.x[0], w.xl) : row_prev(w.x[r], w.x[r > 0 ? r - 1 : 0]);
        const real rgt = r == kWR - 1 ? shift_next(w.x[r], w.xr) : row_next(w.x[r], w.x[r < kWR - 1 ? r + 1 : r]);
        row_extrema<false>(w.x[r], lft, rgt, wave_ballot(u < nt),
                           wave_ballot(p == 0) | wave_ballot(p == q.N - 1), em[0][r], em[1][r]);
#pragma unroll
        for (int ty = 0; ty < 2; ++ty) run[ty] += __popcll(em[ty][r]);
    }
    store_masks(kout.mask + 8 * (int64_t)gt, em, lane);
    if (lane == 0) {
        kout.cnt[2 * gt] = run[0];
        kout.cnt[2 * gt + 1] = run[1];
    }
}

// One wavefront per run of kInitRun consecutive tiles (grid-stride over the
// runs: the grid is capped, so any tile count fits one launch).  The series
// pointers are looked up once per run (again only where the run crosses
// into the next series), and the next tile's samples are requested before
// the current tile's stores, so each tile's loads overlap the previous
// tile's work instead of starting a chain of dependent lookups per tile.
constexpr int kInitRun = 8;

template <typename real, int MODE, bool LOOP>
__global__ __launch_bounds__(kThreads) void k_init_w(SeriesDev sd, InitArgs<real> ia, real* h_out,
                                                     KnotsDev<real> kout)
{
    constexpr int NW = kThreads / kWave;
    const int lane = threadIdx.x & 63, wv = wave_index();
    const int64_t nruns = ((int64_t)sd.ntiles + kInitRun - 1) / kInitRun;
    // one run per wavefront, or (LOOP) a grid-stride loop over the runs
    // where that grid would exceed a launch's 2^32 work-items
    for (int64_t run = (int64_t)blockIdx.x * NW + wv; run < nruns;
         run = LOOP ? run + (int64_t)gridDim.x * NW : nruns) {
        const int g0 = (int)(run * kInitRun);
        const int g1 = min(g0 + kInitRun, sd.ntiles);
        InitSeries<real, MODE> q = init_series<real, MODE>(sd, ia, h_out, sd.tile_series[g0]);
        InitTile<real> cur, nxt;
        init_load<real, MODE>(q, ia, g0, lane, cur);
        for (int gt = g0; gt < g1; ++gt) {
            InitSeries<real, MODE> qn = q;
            if (gt + 1 < g1) {
                if (gt + 1 >= q.tend) qn = init_series<real, MODE>(sd, ia, h_out, sd.tile_series[gt + 1]);
                init_load<real, MODE>(qn, ia, gt + 1, lane, nxt);
            }
            init_store<real, MODE>(q, kout, gt, lane, cur);
            cur = nxt;
            q = qn;
        }
    }
}

// ================================================================ adjoint
// With the extrema sets of every pass frozen (recorded by the forward), one
// sift pass is linear: h' = h - (U + L) / 2.  Its transpose maps the
// gradient mu of h' to that of h: mu - (G_max^T mu + G_min^T mu) / 2, where
// G^T is: per-interval moments of mu (transposed evaluation), transposed
// coefficient formulas, a solve with the transposed spline matrix (its
// interior block is the symmetric interior block of the forward matrix, so
// the forward's windowed solver is reused), transposed right-hand side and
// endpoint fix, and a scatter to the knot samples.

// 1'. Fused per-tile step for 256-sample tiles: one wavefront per tile,
// lane l owns the four consecutive samples 4l .. 4l+3.  Optionally
// mu <- broadcast, then mu <- mu - (G_max^T + G_min^T)/2 for the scatter
// pass, then the moments of mu over the knot intervals of the moment pass:
// each lane sums its own samples run by run; the run a lane ends with is
// continued across lanes by a segmented scan keyed by the run number.
constexpr int kAdjKC = kWT / 2 + 2;

template <typename real>
__device__ __forceinline__ void adj_tile(const AdjTileArgs<real>& a, int gt, int lane, real* delta,
                                         int* tp)
{
    constexpr int T = kWT;
    constexpr int NOKEY = 1 << 28;
    const SeriesDev& sd = a.sd;
    const int s = sd.tile_series[gt];
    const int t = gt - sd.tile0[s];
    const int N = sd.len[s];
    const int t0 = t * T;
    const int nt = min(T, N - t0);
    real* mu = a.mu + sd.off[s];
    const int ub = 4 * lane;

    real v[4];
    bool dirty = false;
    if (a.bcast) {
        const real* src = a.bcast + a.chan_off[sd.chan[s]];
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = ub + q < nt ? src[t0 + ub + q] : (real)0;
        dirty = true;
    } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = ub + q < nt ? mu[t0 + ub + q] : (real)0;
    }
    // ---- scatter of the knot gradients of pass sc: the lane of each knot
    // bit (sample 64 r + lane of the tile) adds the gradient of its knot
    if (a.sc.mask && a.sc.n[2 * s] > 3 && a.sc.n[2 * s + 1] > 3) {
#pragma unroll
        for (int q = 0; q < 4; ++q) delta[ub + q] = (real)0;
        wave_sync();
        for (int ty = 0; ty < 2; ++ty) {
            const int64_t kb = sd.kbase[s] + (ty ? sd.kstride : 0);
            int cb = a.sc.pre[2 * gt + ty];
#pragma unroll
            for (int r = 0; r < kWR; ++r) {
                const uint64_t m = load_scalar(a.sc.mask + 8 * (int64_t)gt + kWR * ty + r);
                if (__builtin_amdgcn_inverse_ballot_w64(m)) {
                    const int u = kWave * r + lane;
                    delta[u] = delta[u] + a.ga[kb + cb + lanes_below(m)];
                }
                cb += __popcll(m);
            }
        }
        wave_sync();
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = v[q] - delta[ub + q] / (real)2;
        dirty = true;
    }
    if (dirty) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (ub + q < nt) mu[t0 + ub + q] = v[q];
    }
    // ---- moments of mu over the intervals of pass mo
    if (!(a.mo.mask && a.mo.n[2 * s] > 3 && a.mo.n[2 * s + 1] > 3)) return;
    for (int ty = 0; ty < 2; ++ty) {
        const int64_t kb = sd.kbase[s] + (ty ? sd.kstride : 0);
        const int cnt = a.mo.cnt[2 * gt + ty], pre = a.mo.pre[2 * gt + ty];
        const int left = pre > 0 ? a.mo.left[2 * gt + ty] : 0;
        // the tile's knot bits (row words) and their positions in rank order
        uint64_t bits[kWR];
#pragma unroll
        for (int r = 0; r < kWR; ++r) bits[r] = load_scalar(a.mo.mask + 8 * (int64_t)gt + kWR * ty + r);
        wave_sync();
        {
            int cb = 0;
#pragma unroll
            for (int r = 0; r < kWR; ++r) {
                if (__builtin_amdgcn_inverse_ballot_w64(bits[r]))
                    tp[cb + lanes_below(bits[r])] = t0 + kWave * r + lane;
                cb += __popcll(bits[r]);
            }
        }
        wave_sync();
        real* e = a.edge + 4 * (4 * (int64_t)gt + 2 * ty);
        if (lane == 0) {
            // empty edge runs (a knot on the tile's first / last sample)
            if (cnt > 0 && (t0 == 0 || tp[0] == t0))
                for (int m = 0; m < 4; ++m) e[m] = (real)0;
            if (cnt > 0 && tp[cnt - 1] == t0 + nt - 1)
                for (int m = 0; m < 4; ++m) e[4 + m] = (real)0;
        }
        // run number of sample u: knots of the tile strictly before it
        // (sample 0 of the series joins run 1, the interval (x_0, x_1])
        // the four samples of a lane lie in knot word ub >> 6: its bits and
        // the knots of the words before it
        const int wd = ub >> 6;
        uint64_t wbits = bits[0];
        int wbefore = 0;
#pragma unroll
        for (int w = 1; w < kWR; ++w) {
            wbefore += wd >= w ? __popcll(bits[w - 1]) : 0;
            wbits = wd == w ? bits[w] : wbits;
        }
        auto key_of = [&](int u) -> int {
            if (u >= nt) return NOKEY;
            if (t0 == 0 && u == 0) return 1;
            const int b = u & 63;
            return wbefore + (b ? __popcll(wbits << (64 - b)) : 0);
        };
        auto dst_of = [&](int r) -> real* {
            if (r >= 1 && r < cnt) return a.mom + 4 * (kb + pre + r - 1);
            return e + (r == 0 ? 0 : 4);
        };
        int key[4];
        real m[4][4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int u = ub + q;
            key[q] = key_of(u);
            real tt = (real)0;
            if (key[q] != NOKEY) {
                const int lp = (t0 == 0 && u == 0) ? 0 : (key[q] >= 1 ? tp[key[q] - 1] : left);
                tt = (real)(t0 + u - lp);
            }
            const real g = key[q] != NOKEY ? v[q] : (real)0;
            m[q][0] = g;
            m[q][1] = g * tt;
            m[q][2] = m[q][1] * tt;
            m[q][3] = m[q][2] * tt;
        }
        // runs inside the lane: the first (maybe continued from the left),
        // complete middle runs, the last (maybe continued to the right)
        const int kfirst = key[0];
        real first[4] = {m[0][0], m[0][1], m[0][2], m[0][3]};
        bool closed = false;
        int kcur = kfirst;
        real cur[4] = {m[0][0], m[0][1], m[0][2], m[0][3]};
#pragma unroll
        for (int q = 1; q < 4; ++q) {
            if (key[q] == kcur) {
#pragma unroll
                for (int z = 0; z < 4; ++z) cur[z] = cur[z] + m[q][z];
            } else {
                if (!closed) {
#pragma unroll
                    for (int z = 0; z < 4; ++z) first[z] = cur[z];
                    closed = true;
                } else if (kcur != NOKEY) {
                    real* d = dst_of(kcur);
#pragma unroll
                    for (int z = 0; z < 4; ++z) d[z] = cur[z];
                }
                kcur = key[q];
#pragma unroll
                for (int z = 0; z < 4; ++z) cur[z] = m[q][z];
            }
        }
        // segmented inclusive scan of the last runs over the lanes (the
        // last-run keys do not decrease with the lane)
        real sc[4] = {cur[0], cur[1], cur[2], cur[3]};
        seg_scan_keys(sc, kcur);
        // lane moves run on every lane
        const int kprev = dpp_prev_i(kcur, NOKEY + 1);
        real prevs[4];
#pragma unroll
        for (int z = 0; z < 4; ++z) prevs[z] = shift_prev(sc[z], (real)0);
        const int knext_first = dpp_next_i(kfirst, NOKEY + 1);
        if (closed && kfirst != NOKEY) {
            // the first run ends inside this lane
            real tot[4];
#pragma unroll
            for (int z = 0; z < 4; ++z)
                tot[z] = (kprev == kfirst) ? prevs[z] + first[z] : first[z];
            real* d = dst_of(kfirst);
#pragma unroll
            for (int z = 0; z < 4; ++z) d[z] = tot[z];
        }
        // the last run ends at this lane's end unless the next lane continues it
        const bool ends = (lane == kWave - 1) || (knext_first != kcur);
        if (ends && kcur != NOKEY) {
            real* d = dst_of(kcur);
#pragma unroll
            for (int z = 0; z < 4; ++z) d[z] = sc[z];
        }
    }
}

// One wavefront per tile (LOOP = false), or a grid-stride loop over the
// tiles where one wave per tile would exceed a launch's 2^32 work-items
// (the wave's LDS slice is reused by its next tile only after a wave
// barrier).  The loop keeps the arguments live across its iterations, which
// spilled scalar registers (20, restored through v_readlane every tile) and
// made the adjoint tile step 44 % slower, so it is used only where it must.
template <typename real, bool LOOP>
__global__ __launch_bounds__(kThreads) void k_adj_tile(AdjTileArgs<real> a)
{
    constexpr int NW = kThreads / kWave;
    __shared__ real s_delta[NW][kWT];
    __shared__ int s_tp[NW][kAdjKC];
    const int lane = threadIdx.x & 63, wv = wave_index();
    if (!LOOP) {
        const int gt = blockIdx.x * NW + wv;
        if (gt < a.sd.ntiles) adj_tile<real>(a, gt, lane, s_delta[wv], s_tp[wv]);
        return;
    }
    for (int gt = blockIdx.x * NW + wv; gt < a.sd.ntiles; gt += gridDim.x * NW) {
        adj_tile<real>(a, gt, lane, s_delta[wv], s_tp[wv]);
        wave_sync();
    }
}

// 2. intervals that leave their tile: the last run of the tile holding the
// left knot plus the first runs of the following tiles up to the right knot
template <typename real>
__global__ void k_adj_span(AdjArgs<real> a)
{
    const int gt = blockIdx.x * blockDim.x + threadIdx.x;
    const SeriesDev& sd = a.sd;
    if (gt >= sd.ntiles) return;
    const int s = sd.tile_series[gt];
    const int n0 = a.ps.n[2 * s], n1 = a.ps.n[2 * s + 1];
    if (!(n0 > 3 && n1 > 3)) return;
    for (int ty = 0; ty < 2; ++ty) {
        const int cnt = a.ps.cnt[2 * gt + ty];
        if (cnt == 0) continue;
        const int g = a.ps.pre[2 * gt + ty] + cnt - 1;
        if (g >= (ty ? n1 : n0) - 1) continue;
        const real* e = a.edge + 4 * (4 * (int64_t)gt + 2 * ty + 1);
        real m0 = e[0], m1 = e[1], m2 = e[2], m3 = e[3];
        // up to the tile holding the next knot (known from the record, so
        // the loads do not wait on each other)
        const int last = sd.tile0[s] + a.ps.right[2 * gt + ty] / kWT;
        int tau = gt + 1;
        // eight tiles at a time: all loads first, then the adds in order
        for (; tau + 7 <= last; tau += 8) {
            real f[8][4];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const real* fp = a.edge + 4 * (4 * (int64_t)(tau + u) + 2 * ty);
#pragma unroll
                for (int m = 0; m < 4; ++m) f[u][m] = fp[m];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                m0 += f[u][0];
                m1 += f[u][1];
                m2 += f[u][2];
                m3 += f[u][3];
            }
        }
        for (; tau <= last; ++tau) {
            const real* f = a.edge + 4 * (4 * (int64_t)tau + 2 * ty);
            m0 += f[0];
            m1 += f[1];
            m2 += f[2];
            m3 += f[3];
        }
        real* dst = a.mom + 4 * (sd.kbase[s] + (ty ? sd.kstride : 0) + g);
        dst[0] = m0;
        dst[1] = m1;
        dst[2] = m2;
        dst[3] = m3;
    }
}

// 3. knot gradients: right-hand side gc of the transposed solve from the
// moments (transposed coefficient formulas, cudaICEEMDAN.cu:394-398), the
// solve, the transposed right-hand side f(a) (:333-341) and endpoint fix
// (:209-291).  One wavefront per window, the forward's window geometry.
// (three waves per SIMD: the compiler is held to 168 registers)
template <typename real, int B, int H>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(3))) void k_adj_knot(AdjArgs<real> a)
{
    constexpr int W = kWave * B;
    constexpr int K = W - 2 * H;
    constexpr int NW = kThreads / kWave;
    __shared__ real s_f[NW][W];            // knot-value gradients of the window rows
    __shared__ int s_pos[NW][W + 2];       // knot positions g0-1 .. g0+W
    __shared__ int s_tpre[NW][kWave];
    __shared__ uint64_t s_tmask[NW][kWave][kTMaskW];
    const int wv = wave_index(), lane = threadIdx.x & 63;
    real* sf = s_f[wv];
    int* sp = s_pos[wv];
    const real third = (real)1 / (real)3;
    // window items and series metadata through the scalar cache (read-only here)
    const unsigned total = load_scalar(a.wincount);
    for (unsigned item = blockIdx.x * NW + wv; item < total; item += gridDim.x * NW) {
        const int s = load_scalar(a.winlist + kWinItem * item);
        const int code = load_scalar(a.winlist + kWinItem * item + 1);
        const int hint = load_scalar(a.winlist + kWinItem * item + 2);
        const int type = code & 1;
        const int w = code >> 1;
        const int n0 = load_scalar(a.ps.n + 2 * s), n1 = load_scalar(a.ps.n + 2 * s + 1);
        if (!(n0 > 3 && n1 > 3)) continue;
        const int n = type ? n1 : n0;
        const int64_t kb = load_scalar(a.sd.kbase + s) + (type ? a.sd.kstride : 0);
        const long g0 = (long)w * K - H;
        // knot positions of the window (and one beyond each side) from the
        // pass's bitmasks: sp index = g - lo
        const long lo = g0 - 1;
        const int glo = (int)max(lo, 0L);
        const int ghi = (int)min(lo + W + 2, (long)n);
        gather_positions<kTMaskW>(a.sd, a.ps.mask, a.ps.pre, a.ps.cnt, s, type, lo, glo, ghi, sp,
                                  s_tpre[wv], s_tmask[wv], hint);
        wave_sync();
        // lane l owns window rows r = l*B + q (knot gb + q); everything a row
        // needs comes from its own lane's registers or a neighbour lane's
        const long gb = g0 + (long)lane * B;
        // positions of knots gb-1 .. gb+B (xk[j] <-> knot gb-1+j), clamped
        // into the series (values past its ends are never used)
        int xk[B + 2];
#pragma unroll
        for (int j = 0; j < B + 2; ++j) {
            const long g = gb - 1 + j;
            xk[j] = sp[min(max(g, 0L), (long)n - 1) - lo];
        }
        // interval gb-1+j (j = 0..B): its moments feed its left knot (row j-1)
        // and its right knot (row j) through the transposed coefficient
        // formulas (cudaICEEMDAN.cu:394-398).  gc = right-hand side of the
        // transposed solve, vp = the part of the knot gradient without the
        // solution; ih = reciprocal interval length.  (The adjoint is compared
        // to tolerance, not bit for bit: one reciprocal per interval.)
        real gc[B], vp[B], ih[B + 1];
        real gr_prev = (real)0, vr_prev = (real)0;
#pragma unroll
        for (int j = 0; j <= B; ++j) {
            const long gg = gb - 1 + j;
            const bool in = gg >= 0 && gg <= n - 2;
            const real* m = a.mom + 4 * (kb + (in ? gg : 0));
            const real m0 = m[0], m1 = m[1], m2 = m[2], m3 = m[3];
            const real h = (real)(xk[j + 1] - xk[j]);
            const real ihh = in ? (real)1 / h : (real)0;
            ih[j] = ihh;
            const real gl = in ? m2 - (real)2 * third * h * m1 - m3 * ihh * third : (real)0;
            const real gr = in ? m3 * ihh * third - h * third * m1 : (real)0;
            const real vl = in ? m0 - m1 * ihh : (real)0;
            const real vr = in ? m1 * ihh : (real)0;
            if (j >= 1) {
                const int q = j - 1;
                const long g = gb + q;
                gc[q] = (g >= 1 && g <= n - 2) ? gl + gr_prev : (real)0;
                vp[q] = vl + vr_prev;
            }
            gr_prev = gr;
            vr_prev = vr;
        }
        real xs[B];
        partition_pcr<real, B>(
            lane,
            [&](int q, real& dl, real& d, real& du, real& f) {
                const long g = gb + q;
                if (g <= 0 || g >= n - 1) {
                    dl = (real)0; d = (real)1; du = (real)0; f = (real)0;
                    return;
                }
                const int hl = xk[q + 1] - xk[q];
                const int hr = xk[q + 2] - xk[q + 1];
                dl = (real)hl;
                du = (real)hr;
                d = (real)(2 * (hl + hr));
                f = gc[q];
            },
            xs);
        // solution of the rows just outside this lane's block
        const real xprev = shift_prev(xs[B - 1], (real)0);
        const real xnext = shift_next(xs[0], (real)0);
        // base knot gradient of every row with both neighbours in the window
#pragma unroll
        for (int q = 0; q < B; ++q) {
            const int r = lane * B + q;
            const long g = gb + q;
            real v = (real)0;
            if (g >= 0 && g < n && r >= 1 && r <= W - 2) {
                const real sm1 = q == 0 ? xprev : xs[q > 0 ? q - 1 : 0];
                const real sp1 = q == B - 1 ? xnext : xs[q < B - 1 ? q + 1 : q];
                v = vp[q];
                if (g - 1 >= 1 && g - 1 <= n - 2) v = v + (real)3 * sm1 * ih[q];
                if (g >= 1 && g <= n - 2) v = v - (real)3 * xs[q] * (ih[q + 1] + ih[q]);
                if (g + 1 >= 1 && g + 1 <= n - 2) v = v + (real)3 * sp1 * ih[q + 1];
            }
            sf[r] = v;
        }
        wave_sync();
        // endpoint fix: a_0 = a_1 (1 + rho) - a_2 rho when it replaced a_0
        if (lane == 0 && n >= 6) {
            const int* xp = sp - lo;   // xp[g] = position of knot g (g in the window)
            const long r0 = 0 - g0;
            if (r0 >= 1 && r0 + 2 <= W - 2 && a.ps.fixf[4 * s + 2 * type]) {
                const real rho = (real)(xp[1] - xp[0]) / (real)(xp[2] - xp[1]);
                sf[r0 + 1] = sf[r0 + 1] + sf[r0] * ((real)1 + rho);
                sf[r0 + 2] = sf[r0 + 2] - sf[r0] * rho;
                sf[r0] = (real)0;
            }
            const long r1 = (long)(n - 1) - g0;
            if (r1 - 2 >= 1 && r1 <= W - 2 && a.ps.fixf[4 * s + 2 * type + 1]) {
                const real rho = (real)(xp[n - 1] - xp[n - 2]) / (real)(xp[n - 2] - xp[n - 3]);
                sf[r1 - 1] = sf[r1 - 1] + sf[r1] * ((real)1 + rho);
                sf[r1 - 2] = sf[r1 - 2] - sf[r1] * rho;
                sf[r1] = (real)0;
            }
        }
        wave_sync();
        for (int r = H + lane; r < H + K; r += kWave) {
            const long g = g0 + r;
            if (g < n) a.ga[kb + g] = sf[r];
        }
        wave_sync();
    }
}

template <typename real>
__global__ void k_adj_seed(const real* gimf, int M, int N, int64_t stride, real* gr)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    gr[p] = gimf[p];
    for (int k = 1; k < M; ++k)
        gr[(int64_t)k * stride + p] = gimf[(int64_t)k * N + p] - gimf[(int64_t)(k - 1) * N + p];
}

// stage level: mu of every realization starts as dL/dr_{k+1} of its channel
// g_x_i = (lambda - mu_i) / I; gacc = sum_i g_x_i (sequential in i)
template <typename real>
__global__ void k_adj_accum(int I, const int* chan_len, const int64_t* chan_off, SeriesDev sd,
                            const real* gr_next, const real* mu, real* gacc)
{
    const int c = blockIdx.y;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= chan_len[c]) return;
    const real lam = gr_next[chan_off[c] + p];
    real acc = (real)0;
    for (int i = 0; i < I; ++i) {
        const int s = c * I + i;
        acc = acc + (lam - mu[sd.off[s] + p]) / (real)I;
    }
    gacc[chan_off[c] + p] = acc;
}

// dots[s] = <g_x_s, E_s> (canonical order), one workgroup per series
template <typename real>
__global__ __launch_bounds__(1024) void k_adj_dot(int I, SeriesDev sd, const int64_t* chan_off,
                                                  const real* gr_next, const real* mu,
                                                  const real* E, real* dots)
{
    __shared__ real sred[1024];
    const int s = blockIdx.x;
    const int N = sd.len[s];
    const real* lam = gr_next + chan_off[sd.chan[s]];
    const real* m = mu + sd.off[s];
    const real* e = E + sd.eoff[s];
    real acc = (real)0;
    for (int p = threadIdx.x; p < N; p += 1024) acc = acc + ((lam[p] - m[p]) / (real)I) * e[p];
    const real d = canon_sum_block<real>(acc, sred);
    if (threadIdx.x == 0) dots[s] = d;
}

// gr_k += gacc (+ eps0 * sum_i dots * (r_k - mean) / ((N-1) std(r_k)), k >= 1)
template <typename real>
__global__ __launch_bounds__(1024) void k_adj_stage(int stage, int I, const int* chan_len,
                                                    const int64_t* chan_off, const real* gacc,
                                                    const real* dots, const real* r_k, real eps0,
                                                    real* gr_k)
{
    __shared__ real sred[1024];
    const int c = blockIdx.x;
    const int N = chan_len[c];
    const int64_t o = chan_off[c];
    real coef = (real)0, mu_r = (real)0;
    if (stage > 0) {
        real acc = (real)0;
        for (int p = threadIdx.x; p < N; p += 1024) acc = acc + r_k[o + p];
        mu_r = canon_sum_block<real>(acc, sred) * (real)(1.0 / (double)N);
        acc = (real)0;
        for (int p = threadIdx.x; p < N; p += 1024) {
            const real dv = r_k[o + p] - mu_r;
            acc = acc + dv * dv;
        }
        const real sq = sqrt(canon_sum_block<real>(acc, sred) * (real)(1.0 / (double)(N - 1)));
        real dsum = (real)0;
        for (int i = 0; i < I; ++i) dsum = dsum + dots[c * I + i];
        coef = eps0 * dsum / ((real)(N - 1) * sq);
    }
    for (int p = threadIdx.x; p < N; p += 1024) {
        real v = gr_k[o + p] + gacc[o + p];
        if (stage > 0) v = v + coef * (r_k[o + p] - mu_r);
        gr_k[o + p] = v;
    }
}

// dL/dy from dL/dx~ with x~ = y / std(y)
template <typename real>
__global__ __launch_bounds__(1024) void k_adj_input(const int* chan_len, const int64_t* chan_off,
                                                    const real* y, const real* gxt, real* gy)
{
    __shared__ real sred[1024];
    const int c = blockIdx.x;
    const int N = chan_len[c];
    const int64_t o = chan_off[c];
    real acc = (real)0;
    for (int p = threadIdx.x; p < N; p += 1024) acc = acc + y[o + p];
    const real mu = canon_sum_block<real>(acc, sred) * (real)(1.0 / (double)N);
    acc = (real)0;
    for (int p = threadIdx.x; p < N; p += 1024) {
        const real dv = y[o + p] - mu;
        acc = acc + dv * dv;
    }
    const real var = canon_sum_block<real>(acc, sred) * (real)(1.0 / (double)(N - 1));
    acc = (real)0;
    for (int p = threadIdx.x; p < N; p += 1024) acc = acc + gxt[o + p] * y[o + p];
    const real dot = canon_sum_block<real>(acc, sred);
    const real sx = (real)1 / sqrt(var);
    for (int p = threadIdx.x; p < N; p += 1024)
        gy[o + p] = sx * gxt[o + p] - sx * sx * sx * dot * (y[o + p] - mu) / (real)(N - 1);
}

// --------------------------------------------------------------------- noise

// value (i, p): Philox4x64-10 block q = p/4, counter (q, i, 0, 0), key (seed, 0);
// u = ((bits >> 11) + 0.5) * 2^-53; Box-Muller on (u0,u1) and (u2,u3).
template <typename real>
__global__ void k_noise(uint64_t seed, int I, int64_t N, real* out, int64_t row_stride)
{
    const int i = blockIdx.y;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (4 * q >= N) return;
    const u64x4 b = philox4x64_10((uint64_t)q, (uint64_t)i, 0, 0, seed, 0);
    const double TWO_PI = 6.283185307179586476925286766559;
    double u[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) u[j] = ((double)(b.v[j] >> 11) + 0.5) * 0x1.0p-53;
    const double r0 = sqrt(-2.0 * log(u[0]));
    const double r1 = sqrt(-2.0 * log(u[2]));
    double z[4];
    z[0] = r0 * cos(TWO_PI * u[1]);
    z[1] = r0 * sin(TWO_PI * u[1]);
    z[2] = r1 * cos(TWO_PI * u[3]);
    z[3] = r1 * sin(TWO_PI * u[3]);
    real* row = out + (int64_t)i * row_stride;
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (4 * q + j < N) row[4 * q + j] = (real)z[j];
}

// ------------------------------------------------------------ stage kernels

// ensemble mean of x_i - h_i over realizations, sequential in i
// (produceFirstIMF :41-55, produceResidue + averageUpdateSignal :582-607)
template <typename real>
__device__ __forceinline__ real acc_stage0(real tem, double invI, real d)
{
    return (real)((double)tem + invI * (double)d);
}

template <typename real>
__global__ void k_ensemble(int stage, int I, const int* chan_len, const int64_t* chan_off,
                           const int64_t* ser_off, const int64_t* ser_eoff, const int* ser_nidx,
                           const real* hA,
                           const real* hB, const int* mode_parity, int fixed_parity, const real* E,
                           const double* inv, const double* sq, real eps0, const real* xt,
                           real* r, real* imf_base, const int64_t* chan_imf_off)
{
    using P = typename Promote<real>::type;
    const int c = blockIdx.y;
    const int N = chan_len[c];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const int64_t co = chan_off[c];
    // row `stage` of channel c's [M][N] output
    real* imf_k = imf_base + chan_imf_off[c] + (int64_t)stage * N;
    const int s0 = c * I;
    real tem = (real)0;
    if (stage == 0) {
        const double invI = 1.0 / (double)I;
        const real xv = xt[co + p];
        for (int i = 0; i < I; ++i) {
            const int s = s0 + i;
            const int par = mode_parity ? (mode_parity[s] & 1) : fixed_parity;
            const real hv = (par ? hB : hA)[ser_off[s] + p];
            const real e = E[ser_eoff[s] + p];
            const real x = (real)((P)xv + (P)(eps0 * e) * (P)inv[ser_nidx[s]]);
            tem = acc_stage0<real>(tem, invI, x - hv);
        }
        r[co + p] = tem;
        imf_k[p] = xv - tem;
    } else {
        const real rv = r[co + p];
        const real sqc = (real)sq[c];
        for (int i = 0; i < I; ++i) {
            const int s = s0 + i;
            const int par = mode_parity ? (mode_parity[s] & 1) : fixed_parity;
            const real hv = (par ? hB : hA)[ser_off[s] + p];
            const real e = E[ser_eoff[s] + p];
            const real x = rv + (eps0 * e) * sqc;
            tem = tem + (x - hv);
        }
        imf_k[p] = rv - tem / (real)I;
        r[co + p] = tem / (real)I;
    }
}

// x~ = y * (1 / sqrt(var(y)))  (standardizeRunning, cudaICEEMDAN.cu:57-66)
template <typename real>
__global__ void k_standardize(const real* y, const real* var, real* xt, int64_t N)
{
    using P = typename Promote<real>::type;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const P inv = 1.0 / (P)sqrt(var[0]);
    xt[p] = (real)((P)y[p] * inv);
}

// out = 1/sqrt(var) (take_sqrt = 0) or sqrt(var) (take_sqrt = 1), as double
template <typename real>
__global__ void k_inv_sqrt(const real* var, double* out, int n, int take_sqrt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const real sd = sqrt(var[i]);
    out[i] = take_sqrt ? (double)sd : 1.0 / (double)sd;
}

// elementwise kernels: grid-stride with 64-bit indices (capped grids)
template <typename real>
__global__ void k_sub(real* w, const real* e, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        w[i] = w[i] - e[i];
}

template <typename real>
__global__ void k_copy(real* dst, const real* src, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// dst[s] = (parity[s] odd ? hB : hA)[s] for every series (RILLING export)
template <typename real>
__global__ void k_select_copy(SeriesDev sd, const real* hA, const real* hB, const int* parity,
                              real* dst)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    for (int s = blockIdx.y; s < sd.nseries; s += gridDim.y) {
        if (p >= sd.len[s]) continue;
        const int64_t o = sd.off[s] + p;
        dst[o] = ((parity[s] & 1) ? hB : hA)[o];
    }
}

// left/right[2 t + ty] = positions of the last knot before tile t and of the
// first knot after it (from the prefix's neighbour knots), for the
// adjoint's pass records
__global__ void k_rec_left(const int* nbx, int* left, int* right, int ntiles)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 2 * ntiles) {
        left[i] = nbx[6 * (int64_t)(i >> 1) + 3 * (i & 1)];
        right[i] = nbx[6 * (int64_t)(i >> 1) + 3 * (i & 1) + 2];
    }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Order-independent 64-bit digest of the knot sets of a pass input, per
// series: the sum (mod 2^64) over the series' tiles of a mix of each of the
// tile's eight bitmask words with its (tile, type, row) index.  Two passes
// with equal digests met the same knot sets (up to a hash collision).
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    // splitmix64 finalizer (Steele, Lea, Flood 2014)
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void k_knot_digest(const int* tile_series, const int* tile0, const uint64_t* mask,
                              int ntiles, unsigned long long* out)
{
    const int gt = blockIdx.x * blockDim.x + threadIdx.x;
    if (gt >= ntiles) return;
    const int s = tile_series[gt];
    const uint64_t t = (uint64_t)(gt - tile0[s]);
    uint64_t h = 0;
#pragma unroll
    for (int q = 0; q < 2 * kWR; ++q) h += mix64(mask[8 * (int64_t)gt + q] ^ mix64(8 * t + q + 1));
    atomicAdd(out + s, (unsigned long long)h);
}

__global__ void k_check_index(const int* x, int64_t N, int* err)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N && x[i] != (int)i) atomicOr(err, ICM_DEVERR_INDEX);
}

// ---------------------------------------------------------------- launchers

static inline unsigned blocks_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }
// grid of a grid-stride launch: at most 2^20 workgroups (2^28 work-items)
static inline unsigned capped_blocks(int64_t n, int b)
{
    const int64_t g = (n + b - 1) / b;
    return (unsigned)(g < 1 ? 1 : (g > (1 << 20) ? (1 << 20) : g));
}
// workgroups of kThreads for n items, b per workgroup, if that grid fits one
// launch (fewer than 2^32 work-items); else 0 (the grid-stride instance)
static inline unsigned full_blocks(int64_t n, int b)
{
    const int64_t g = std::max<int64_t>(1, (n + b - 1) / b);
    return g * kThreads < ((int64_t)1 << 32) ? (unsigned)g : 0u;
}

// Workgroups of `kernel` (kThreads each) that the device keeps resident at
// once, per device (the grid of a persistent launch).
template <typename K>
static int resident_groups(K kernel)
{
    static std::mutex mu;
    static std::map<std::pair<const void*, int>, int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    const auto key = std::make_pair((const void*)kernel, dev);
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int cus = 0, per = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, kThreads, 0) != hipSuccess ||
        per <= 0)
        per = 4;
    cache[key] = cus * per;
    return cus * per;
}

template <typename real> struct Geom;
template <> struct Geom<double> { static constexpr int B = 8, H = 32; };
template <> struct Geom<float> { static constexpr int B = 8, H = 16; };

// tile size (samples per tile) is a plan parameter: sd.cap = T/2 + 2
template <typename real, int MODE>
static void launch_init_mode(bool loop, dim3 grid, hipStream_t st, const SeriesDev& sd,
                             const InitArgs<real>& ia, real* h_out, const KnotsDev<real>& kout)
{
    if (loop)
        hipLaunchKernelGGL((k_init_w<real, MODE, true>), grid, dim3(kThreads), 0, st, sd, ia, h_out,
                           kout);
    else
        hipLaunchKernelGGL((k_init_w<real, MODE, false>), grid, dim3(kThreads), 0, st, sd, ia, h_out,
                           kout);
}

template <typename real>
hipError_t launch_init(const SeriesDev& sd, const InitArgs<real>& ia, real* h_out,
                       const KnotsDev<real>& kout, const PassCtl&, hipStream_t st, bool force_loop)
{
    const int64_t runs = ((int64_t)sd.ntiles + kInitRun - 1) / kInitRun;
    const unsigned full = force_loop ? 0u : full_blocks(runs, kThreads / kWave);
    const bool loop = full == 0;
    const dim3 grid(loop ? capped_blocks(runs, kThreads / kWave) : full);
    switch (ia.mode) {
    case INIT_COPY: launch_init_mode<real, INIT_COPY>(loop, grid, st, sd, ia, h_out, kout); break;
    case INIT_RESID: launch_init_mode<real, INIT_RESID>(loop, grid, st, sd, ia, h_out, kout); break;
    case INIT_STAGE0: launch_init_mode<real, INIT_STAGE0>(loop, grid, st, sd, ia, h_out, kout); break;
    case INIT_KNOTS: launch_init_mode<real, INIT_KNOTS>(loop, grid, st, sd, ia, h_out, kout); break;
    default: launch_init_mode<real, INIT_STAGEK>(loop, grid, st, sd, ia, h_out, kout); break;
    }
    return hipGetLastError();
}

template <typename real>
hipError_t launch_knot_prefix(const SeriesDev& sd, const KnotsDev<real>& k, const real* h,
                              const PassCtl& pc, const int* active, hipStream_t st)
{
    if (sd.max_tiles <= kPWTiles)
        hipLaunchKernelGGL((k_knot_prefix_w<real>), dim3(blocks_for(sd.nseries, kThreads / kWave)),
                           dim3(kThreads), 0, st, sd, k, h, pc, active);
    else if (sd.max_tiles <= kPT)
        hipLaunchKernelGGL((k_knot_prefix_1<real>), dim3(sd.nseries), dim3(kPT), 0, st, sd, k, h,
                           pc, active);
    else
        hipLaunchKernelGGL((k_knot_prefix_b<real>), dim3(sd.nseries), dim3(kPT), 0, st, sd, k, h,
                           pc, active);
    return hipGetLastError();
}

// grid of a k_sift_w launch: with a work queue, what the device keeps
// resident (no more than one wave per segment); without, one wave per segment
template <typename K>
static dim3 sift_grid(K kernel, const unsigned* queue, int nsegs)
{
    const int64_t need = blocks_for(nsegs, kThreads / kWave);
    if (!queue) return dim3((unsigned)need);
    return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(need, resident_groups(kernel))));
}

template <typename real, bool RILL, int KC>
static void launch_sift_w(const SiftArgs<real>& a, int nsegs, hipStream_t st)
{
    if (a.queue) {
        auto k = k_sift_w<real, RILL, KC, true>;
        hipLaunchKernelGGL(k, sift_grid(k, a.queue, nsegs), dim3(kThreads), 0, st, a, nsegs);
    } else {
        auto k = k_sift_w<real, RILL, KC, false>;
        hipLaunchKernelGGL(k, sift_grid(k, a.queue, nsegs), dim3(kThreads), 0, st, a, nsegs);
    }
}

template <typename real>
hipError_t launch_sift(const SiftArgs<real>& a, int nsegs, hipStream_t st)
{
    if (!a.ovf_segs) {
        if (a.sx_cnt) launch_sift_w<real, true, kWKC>(a, nsegs, st);
        else launch_sift_w<real, false, kWKC>(a, nsegs, st);
        return hipGetLastError();
    }
    // sparse pass: if few tiles of the input overflow the small knot tables
    // (the knot prefix counted them), small tables (more waves per SIMD) and
    // then the listed overflow tiles with full tables; else the small-table
    // launch returns at once and the second launch sifts every segment.
    // Its grid is what the full-table kernel keeps resident (four
    // workgroups per CU), so it costs little when it has nothing to do.
    const dim3 lgrid(std::min<unsigned>(blocks_for(nsegs, kThreads / kWave), 1024u));
    if (a.sx_cnt) {
        launch_sift_w<real, true, kWKS>(a, nsegs, st);
        hipLaunchKernelGGL((k_sift_list<real, true>), lgrid, dim3(kThreads), 0, st, a, nsegs);
    } else {
        launch_sift_w<real, false, kWKS>(a, nsegs, st);
        hipLaunchKernelGGL((k_sift_list<real, false>), lgrid, dim3(kThreads), 0, st, a, nsegs);
    }
    return hipGetLastError();
}

template <typename real>
hipError_t launch_solve(const SolveArgs<real>& a, int grid, hipStream_t st)
{
    if (a.nak)
        hipLaunchKernelGGL((k_solve<real, Geom<real>::B, Geom<real>::H, true>), dim3(grid),
                           dim3(kThreads), 0, st, a);
    else
        hipLaunchKernelGGL((k_solve<real, Geom<real>::B, Geom<real>::H, false>), dim3(grid),
                           dim3(kThreads), 0, st, a);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_rilling_decide(const SeriesDev& sd, const int* nk, const int* sx_cnt,
                                 const int* sx_mark, int* active, int* stop_pass, int* subs,
                                 int pass, int cap, const unsigned* run, unsigned* n_active,
                                 hipStream_t st)
{
    hipLaunchKernelGGL((k_rilling_decide<real>), dim3(sd.nseries), dim3(kWave), 0, st, sd, nk,
                       sx_cnt, sx_mark, active, stop_pass, subs, pass, cap, run, n_active);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_series_var(const real* base, const int64_t* off, int64_t stride, const int* len,
                             int nser, real* var_out, hipStream_t st)
{
    hipLaunchKernelGGL((k_series_var<real>), dim3(nser), dim3(1024), 0, st, base, off, stride, len,
                       var_out);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_noise(uint64_t seed, int I, int64_t N, real* out, int64_t row_stride,
                        hipStream_t st)
{
    const int64_t nq = (N + 3) / 4;
    hipLaunchKernelGGL((k_noise<real>), dim3(blocks_for(nq, 256), I), dim3(256), 0, st, seed, I, N,
                       out, row_stride);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_ensemble(int stage, int nchan, int I, int max_len, const int* chan_len,
                           const int64_t* chan_off, const int64_t* ser_off,
                           const int64_t* ser_eoff, const int* ser_nidx, const real* hA, const real* hB,
                           const int* mode_parity, int fixed_parity, const real* E,
                           const double* inv, const double* sq, real eps0, const real* xt,
                           real* r, real* imf_base, const int64_t* chan_imf_off, hipStream_t st)
{
    hipLaunchKernelGGL((k_ensemble<real>), dim3(blocks_for(max_len, 256), nchan), dim3(256), 0, st,
                       stage, I, chan_len, chan_off, ser_off, ser_eoff, ser_nidx, hA, hB, mode_parity,
                       fixed_parity, E, inv, sq, eps0, xt, r, imf_base, chan_imf_off);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_standardize(const real* y, const real* var, real* xt, int64_t N, hipStream_t st)
{
    hipLaunchKernelGGL((k_standardize<real>), dim3(blocks_for(N, 256)), dim3(256), 0, st, y, var,
                       xt, N);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_inv_sqrt(const real* var, double* out, int n, int take_sqrt, hipStream_t st)
{
    hipLaunchKernelGGL((k_inv_sqrt<real>), dim3(blocks_for(n, 256)), dim3(256), 0, st, var, out, n,
                       take_sqrt);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_sub(real* w, const real* e, int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL((k_sub<real>), dim3(capped_blocks(n, 256)), dim3(256), 0, st, w, e, n);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_copy(real* dst, const real* src, int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL((k_copy<real>), dim3(capped_blocks(n, 256)), dim3(256), 0, st, dst, src, n);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_select_copy(const SeriesDev& sd, int max_len, const real* hA, const real* hB,
                              const int* parity, real* dst, hipStream_t st)
{
    const unsigned gy = (unsigned)(sd.nseries < 65535 ? sd.nseries : 65535);
    hipLaunchKernelGGL((k_select_copy<real>), dim3(blocks_for(max_len, 256), gy), dim3(256), 0, st,
                       sd, hA, hB, parity, dst);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_adj_tile(const AdjTileArgs<real>& a, hipStream_t st, bool force_loop)
{
    if (2 * (a.sd.cap - 2) != kWT) return hipErrorInvalidValue;
    if (const unsigned full = force_loop ? 0u : full_blocks(a.sd.ntiles, kThreads / kWave))
        hipLaunchKernelGGL((k_adj_tile<real, false>), dim3(full), dim3(kThreads), 0, st, a);
    else
        hipLaunchKernelGGL((k_adj_tile<real, true>), dim3(capped_blocks(a.sd.ntiles, kThreads / kWave)),
                           dim3(kThreads), 0, st, a);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_adj_span_knot(const AdjArgs<real>& a, int solve_grid, hipStream_t st)
{
    hipLaunchKernelGGL((k_adj_span<real>), dim3(blocks_for(a.sd.ntiles, 256)), dim3(256), 0, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_adj_knot<real, Geom<real>::B, Geom<real>::H>), dim3(solve_grid),
                       dim3(kThreads), 0, st, a);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_adj_seed(const real* gimf, int M, int64_t N, int64_t stride, real* gr,
                           hipStream_t st)
{
    hipLaunchKernelGGL((k_adj_seed<real>), dim3(blocks_for(N, 256)), dim3(256), 0, st, gimf, M,
                       (int)N, stride, gr);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_adj_stage(int stage, int nchan, int I, int max_len, const int* chan_len,
                            const int64_t* chan_off, const SeriesDev& sd, const real* gr_next,
                            const real* mu, const real* E, const real* r_k, real eps0, real* gr_k,
                            real* gacc, real* dots, hipStream_t st)
{
    hipLaunchKernelGGL((k_adj_accum<real>), dim3(blocks_for(max_len, 256), nchan), dim3(256), 0, st,
                       I, chan_len, chan_off, sd, gr_next, mu, gacc);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_adj_dot<real>), dim3(sd.nseries), dim3(1024), 0, st, I, sd, chan_off,
                       gr_next, mu, E, dots);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL((k_adj_stage<real>), dim3(nchan), dim3(1024), 0, st, stage, I, chan_len,
                       chan_off, gacc, dots, r_k, eps0, gr_k);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_adj_input(int nchan, const int* chan_len, const int64_t* chan_off, const real* y,
                            const real* gxt, real* gy, hipStream_t st)
{
    hipLaunchKernelGGL((k_adj_input<real>), dim3(nchan), dim3(1024), 0, st, chan_len, chan_off, y,
                       gxt, gy);
    return hipGetLastError();
}

hipError_t launch_rec_left(const int* nbx, int* left, int* right, int ntiles, hipStream_t st)
{
    hipLaunchKernelGGL(k_rec_left, dim3(blocks_for(2 * (int64_t)ntiles, 256)), dim3(256), 0, st, nbx,
                       left, right, ntiles);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_u32, dim3(capped_blocks(n, 256)), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_knot_digest(const int* tile_series, const int* tile0, const uint64_t* mask,
                              int ntiles, uint64_t* out, int nseries, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(2 * (int64_t)nseries, 256)), dim3(256), 0, st,
                       (uint32_t*)out, 0u, 2 * (int64_t)nseries);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_knot_digest, dim3(blocks_for(ntiles, 256)), dim3(256), 0, st, tile_series,
                       tile0, mask, ntiles, (unsigned long long*)out);
    return hipGetLastError();
}

hipError_t launch_check_index(const int* d_x, int64_t N, int* err, hipStream_t st)
{
    hipLaunchKernelGGL(k_check_index, dim3(blocks_for(N, 256)), dim3(256), 0, st, d_x, N, err);
    return hipGetLastError();
}

#define ICM_INSTANTIATE(real)                                                                      \
    template hipError_t launch_init<real>(const SeriesDev&, const InitArgs<real>&, real*,          \
                                          const KnotsDev<real>&, const PassCtl&, hipStream_t,      \
                                          bool);                                                   \
    template hipError_t launch_sift<real>(const SiftArgs<real>&, int, hipStream_t);                \
    template hipError_t launch_knot_prefix<real>(const SeriesDev&, const KnotsDev<real>&,          \
                                                 const real*, const PassCtl&, const int*,          \
                                                 hipStream_t);                                     \
    template hipError_t launch_solve<real>(const SolveArgs<real>&, int, hipStream_t);              \
    template hipError_t launch_rilling_decide<real>(const SeriesDev&, const int*, const int*,      \
                                                    const int*, int*, int*, int*, int, int,        \
                                                    const unsigned*, unsigned*, hipStream_t);      \
    template hipError_t launch_series_var<real>(const real*, const int64_t*, int64_t, const int*,  \
                                                int, real*, hipStream_t);                          \
    template hipError_t launch_noise<real>(uint64_t, int, int64_t, real*, int64_t, hipStream_t);  \
    template hipError_t launch_ensemble<real>(int, int, int, int, const int*, const int64_t*,      \
                                              const int64_t*, const int64_t*, const int*,          \
                                              const real*,                                         \
                                              const real*, const int*, int, const real*,           \
                                              const double*, const double*, real, const real*,     \
                                              real*, real*, const int64_t*, hipStream_t);          \
    template hipError_t launch_standardize<real>(const real*, const real*, real*, int64_t,         \
                                                 hipStream_t);                                     \
    template hipError_t launch_inv_sqrt<real>(const real*, double*, int, int, hipStream_t);        \
    template hipError_t launch_sub<real>(real*, const real*, int64_t, hipStream_t);                \
    template hipError_t launch_copy<real>(real*, const real*, int64_t, hipStream_t);               \
    template hipError_t launch_select_copy<real>(const SeriesDev&, int, const real*, const real*,  \
                                                 const int*, real*, hipStream_t);

#define ICM_INSTANTIATE_ADJ(real)                                                                  \
    template hipError_t launch_adj_tile<real>(const AdjTileArgs<real>&, hipStream_t, bool);        \
    template hipError_t launch_adj_span_knot<real>(const AdjArgs<real>&, int, hipStream_t);        \
    template hipError_t launch_adj_seed<real>(const real*, int, int64_t, int64_t, real*,           \
                                              hipStream_t);                                        \
    template hipError_t launch_adj_stage<real>(int, int, int, int, const int*, const int64_t*,     \
                                               const SeriesDev&, const real*, const real*,         \
                                               const real*, const real*, real, real*, real*,       \
                                               real*, hipStream_t);                                \
    template hipError_t launch_adj_input<real>(int, const int*, const int64_t*, const real*,       \
                                               const real*, real*, hipStream_t);

ICM_INSTANTIATE(float)
ICM_INSTANTIATE(double)
ICM_INSTANTIATE_ADJ(float)
ICM_INSTANTIATE_ADJ(double)

}  // namespace icm
