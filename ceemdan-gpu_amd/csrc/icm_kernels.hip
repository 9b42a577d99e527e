// HIP kernels of the ICEEMDAN hot path (gfx950, wave64).
//
// One sift pass of every series is two launches:
//   k_solve  one wavefront per window of 64*B knots: natural-spline system
//            rows, per-lane elimination of B rows, parallel cyclic reduction
//            across the 64 lanes, back substitution -> c of the kept rows
//   k_sift   one workgroup per tile of T samples: interval of every sample
//            from the tile's extrema flags, interval coefficients in LDS,
//            upper/lower envelope, mean, subtract; then the extrema of the new
//            samples are compacted into the next pass's knot arrays with a
//            single-pass decoupled look-back over the tiles of the series.
// Reference counterparts: find_extrema_shfl_* :92-148, CUB ExclusiveSum
// :859-883, select_extrema_* :150-207, setBoundary* :209-291,
// tridiagonal_setup :310-343, gtsv2_nopivot :899-921, spline_coefficients
// :382-401, interpolate :403-439, averageUppperLower :441-454,
// produceSX/thresholdJudge :456-492, updateRealizations :527-544
// (all cudaICEEMDAN.cu).
#include "icm_device.h"
#include "icm_kernels.h"
#include "icm_errors.h"

namespace icm {

// promotion type used where the reference mixes float data with double
// literals (cudaICEEMDAN.cu:50, :64, :76); identity for double
template <typename real> struct Promote { using type = double; };

// ------------------------------------------------------------ tile helpers

template <typename real, int T>
struct TileGeom {
    int gt, s, t, N, t0, nt;
};

template <int T>
__device__ __forceinline__ int take_ticket(unsigned* tickets, uint32_t epoch, int* s_slot)
{
    if (threadIdx.x == 0) *s_slot = (int)atomicAdd(&tickets[epoch], 1u);
    __syncthreads();
    return *s_slot;
}

// Extrema flags over the tile samples u = p - t0 in [0, nt), from LDS
// samples sh[u + 2] (sh covers p in [t0-2, t0+nt]).  mask[type*NCH + chunk].
template <typename real, int T>
__device__ __forceinline__ void tile_flags(const real* sh, int N, int t0, int nt, uint64_t* mask)
{
    constexpr int NCH = T / 64;
    const int tid = threadIdx.x, lane = tid & 63;
#pragma unroll
    for (int r = 0; r < T / kThreads; ++r) {
        const int u = r * kThreads + tid;
        const int p = t0 + u;
        bool fmx = false, fmn = false;
        if (u < nt) {
            if (p == 0 || p == N - 1) {
                fmx = true;
                fmn = true;
            } else {
                const real v = sh[u + 2], lft = sh[u + 1], rgt = sh[u + 3];
                fmx = (v > lft) && (v > rgt);
                fmn = (v < lft) && (v < rgt);
            }
        }
        const uint64_t bm = wave_ballot(fmx);
        const uint64_t bn = wave_ballot(fmn);
        if (lane == 0) {
            mask[u >> 6] = bm;
            mask[NCH + (u >> 6)] = bn;
        }
    }
}

// chunk exclusive prefixes: pre[type*(NCH+1) + chunk], total at [.. + NCH]
template <int T>
__device__ __forceinline__ void chunk_prefix(const uint64_t* mask, int* pre)
{
    constexpr int NCH = T / 64;
    const int tid = threadIdx.x;
    if (tid < 2) {
        int run = 0;
        for (int ch = 0; ch < NCH; ++ch) {
            pre[tid * (NCH + 1) + ch] = run;
            run += __popcll(mask[tid * NCH + ch]);
        }
        pre[tid * (NCH + 1) + NCH] = run;
    }
}

template <int T>
__device__ __forceinline__ int rank_in_tile(const uint64_t* mask, const int* pre, int type, int u)
{
    constexpr int NCH = T / 64;
    const int ch = u >> 6;
    const uint64_t m = mask[type * NCH + ch];
    const uint64_t below = (u & 63) ? (m & ((1ull << (u & 63)) - 1ull)) : 0ull;
    return pre[type * (NCH + 1) + ch] + __popcll(below);
}

// Compact the extrema of the new samples (sh over [t0-2, t0+nt], new values
// valid on [t0-1, t0+nt]) into the knot arrays of the next pass.
template <typename real, int T>
__device__ void emit_tile(const SeriesDev& sd, const KnotsDev<real>& kout, const PassCtl& pc,
                          const TileGeom<real, T>& g, const real* sh, uint64_t* mask, int* cpre,
                          int* s_pre)
{
    constexpr int NCH = T / 64;
    const int tid = threadIdx.x;
    tile_flags<real, T>(sh, g.N, g.t0, g.nt, mask);
    __syncthreads();
    chunk_prefix<T>(mask, cpre);
    __syncthreads();
    const uint32_t cmx = (uint32_t)cpre[NCH];
    const uint32_t cmn = (uint32_t)cpre[(NCH + 1) + NCH];
    if (tid == 0) {
        const bool first = (g.t == 0);
        status_store(&pc.status[g.gt],
                     pack_status(pc.epoch, first ? kFlagInclusive : kFlagAggregate, cmx, cmn));
        uint32_t pmx = 0, pmn = 0;
        if (!first) {
            int q = g.gt - 1;
            const int qfirst = g.gt - g.t;
            uint32_t spins = 0;
            while (true) {
                const uint64_t w = status_load(&pc.status[q]);
                if (st_epoch(w) != (pc.epoch & 0xFFFFu) || st_flag(w) == 0u) {
                    if (++spins > kSpinCap) {
                        atomicOr(pc.err, ICM_DEVERR_SPIN);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                pmx += st_cmax(w);
                pmn += st_cmin(w);
                if (st_flag(w) == kFlagInclusive || q == qfirst) break;
                --q;
            }
            status_store(&pc.status[g.gt],
                         pack_status(pc.epoch, kFlagInclusive, pmx + cmx, pmn + cmn));
        }
        s_pre[0] = (int)pmx;
        s_pre[1] = (int)pmn;
    }
    __syncthreads();
    const int pmx = s_pre[0], pmn = s_pre[1];
    const int64_t kb = sd.kbase[g.s];
#pragma unroll
    for (int r = 0; r < T / kThreads; ++r) {
        const int u = r * kThreads + tid;
        if (u < g.nt) {
            const int ch = u >> 6;
            const int bit = u & 63;
            const real v = sh[u + 2];
            if ((mask[ch] >> bit) & 1ull) {
                const int64_t o = kb + pmx + rank_in_tile<T>(mask, cpre, 0, u);
                kout.x[o] = g.t0 + u;
                kout.y[o] = v;
            }
            if ((mask[NCH + ch] >> bit) & 1ull) {
                const int64_t o = sd.kstride + kb + pmn + rank_in_tile<T>(mask, cpre, 1, u);
                kout.x[o] = g.t0 + u;
                kout.y[o] = v;
            }
        }
    }
    if (tid == 0) {
        kout.pre[2 * g.gt] = pmx;
        kout.pre[2 * g.gt + 1] = pmn;
        if (g.t == sd.ntl[g.s] - 1) {
            const int nmx = pmx + (int)cmx, nmn = pmn + (int)cmn;
            kout.n[2 * g.s] = nmx;
            kout.n[2 * g.s + 1] = nmn;
            const int wmx = (nmx + pc.K - 1) / pc.K;
            const int wmn = (nmn + pc.K - 1) / pc.K;
            const unsigned base = atomicAdd(&pc.wincount[pc.epoch], (unsigned)(wmx + wmn));
            for (int w = 0; w < wmx; ++w) {
                pc.winlist[2 * (base + w)] = g.s;
                pc.winlist[2 * (base + w) + 1] = (w << 1);
            }
            for (int w = 0; w < wmn; ++w) {
                pc.winlist[2 * (base + wmx + w)] = g.s;
                pc.winlist[2 * (base + wmx + w) + 1] = (w << 1) | 1;
            }
        }
    }
}

template <typename real, int T>
__device__ __forceinline__ TileGeom<real, T> tile_geom(const SeriesDev& sd, int gt)
{
    TileGeom<real, T> g;
    g.gt = gt;
    g.s = sd.tile_series[gt];
    g.t = gt - sd.tile0[g.s];
    g.N = sd.len[g.s];
    g.t0 = g.t * T;
    g.nt = min(T, g.N - g.t0);
    return g;
}

// ------------------------------------------------------------- stage init

template <typename real>
__device__ __forceinline__ real init_value(const InitArgs<real>& ia, const SeriesDev& sd, int s,
                                           int p)
{
    using P = typename Promote<real>::type;
    if (ia.mode == INIT_COPY) return ia.src[sd.off[s] + p];
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

template <typename real, int T>
__global__ __launch_bounds__(kThreads) void k_init(SeriesDev sd, InitArgs<real> ia, real* h_out,
                                                   KnotsDev<real> kout, PassCtl pc)
{
    constexpr int NCH = T / 64;
    __shared__ real sh[T + 3];
    __shared__ uint64_t mask[2 * NCH];
    __shared__ int cpre[2 * (NCH + 1)];
    __shared__ int s_misc[4];
    const int gt = take_ticket<T>(pc.tickets, pc.epoch, &s_misc[0]);
    const TileGeom<real, T> g = tile_geom<real, T>(sd, gt);
    real* out = h_out + sd.off[g.s];
    for (int v = threadIdx.x; v < g.nt + 2; v += kThreads) {
        const int p = g.t0 - 1 + v;
        if (p >= 0 && p < g.N) {
            const real x = init_value<real>(ia, sd, g.s, p);
            sh[v + 1] = x;
            if (v >= 1 && v <= g.nt) out[p] = x;
        }
    }
    __syncthreads();
    emit_tile<real, T>(sd, kout, pc, g, sh, mask, cpre, &s_misc[2]);
}

// ---------------------------------------------------------------- sift pass

template <typename real, int T>
__global__ __launch_bounds__(kThreads) void k_sift(SiftArgs<real> a)
{
    constexpr int NCH = T / 64;
    constexpr int KC = T / 2 + 8;
    __shared__ real sh[T + 3];
    __shared__ uint64_t mask[2 * NCH];
    __shared__ int cpre[2 * (NCH + 1)];
    __shared__ int skp[2][KC];
    __shared__ real ska[2][KC], skc[2][KC], skb[2][KC], skd[2][KC];
    __shared__ int s_misc[8];
    __shared__ int s_red[2 * (kThreads / kWave)];

    const int tid = threadIdx.x;
    const SeriesDev& sd = a.sd;
    const int gt = take_ticket<T>(a.pc.tickets, a.pc.epoch, &s_misc[0]);
    const TileGeom<real, T> g = tile_geom<real, T>(sd, gt);
    if (a.active && !a.active[g.s]) return;

    const real* h = a.h_in + sd.off[g.s];
    real* hout = a.h_out + sd.off[g.s];
    for (int q = tid; q < g.nt + 3; q += kThreads) {
        const int p = g.t0 - 2 + q;
        sh[q] = (p >= 0 && p < g.N) ? h[p] : (real)0;
    }
    __syncthreads();
    tile_flags<real, T>(sh, g.N, g.t0, g.nt, mask);
    if (tid == 0) {
        // extrema flags of sample t0-1 (needed for its interval)
        int fmx = 0, fmn = 0;
        const int p = g.t0 - 1;
        if (p == 0) {
            fmx = fmn = 1;
        } else if (p > 0) {
            fmx = (sh[1] > sh[0]) && (sh[1] > sh[2]);
            fmn = (sh[1] < sh[0]) && (sh[1] < sh[2]);
        }
        s_misc[2] = fmx;
        s_misc[3] = fmn;
    }
    __syncthreads();
    chunk_prefix<T>(mask, cpre);
    __syncthreads();

    const int pre0 = a.kin.pre[2 * gt], pre1 = a.kin.pre[2 * gt + 1];
    const int n0 = a.kin.n[2 * g.s], n1 = a.kin.n[2 * g.s + 1];
    const int fprev0 = s_misc[2], fprev1 = s_misc[3];
    // interval k of sample p: pos[k] < p <= pos[k+1], p = 0 -> 0 (cudaICEEMDAN.cu:418-433)
    auto interval_of = [&](int type, int p) -> int {
        if (p == 0) return 0;
        const int pre = type ? pre1 : pre0;
        int cl;
        if (p == g.t0 - 1) {
            cl = pre - (type ? fprev1 : fprev0);
        } else {
            const int u = p - g.t0;
            if (u >= g.nt) cl = pre + cpre[type * (NCH + 1) + NCH];
            else cl = pre + rank_in_tile<T>(mask, cpre, type, u);
        }
        return cl - 1;
    };
    const int p_first = max(g.t0 - 1, 0);
    const int p_last = min(g.t0 + g.nt, g.N - 1);
    const int klo0 = interval_of(0, p_first), khi0 = interval_of(0, p_last) + 1;
    const int klo1 = interval_of(1, p_first), khi1 = interval_of(1, p_last) + 1;
    int cnt0 = khi0 - klo0 + 1, cnt1 = khi1 - klo1 + 1;
    if (cnt0 > KC || cnt1 > KC || cnt0 < 2 || cnt1 < 2 || khi0 >= n0 || khi1 >= n1) {
        if (tid == 0) atomicOr(a.pc.err, ICM_DEVERR_KNOTCAP);
        return;
    }
    const int64_t kb = sd.kbase[g.s];
    for (int i = tid; i < cnt0 + cnt1; i += kThreads) {
        const int type = (i >= cnt0);
        const int idx = type ? i - cnt0 : i;
        const int64_t gi = kb + (type ? sd.kstride : 0) + (type ? klo1 : klo0) + idx;
        skp[type][idx] = a.kin.x[gi];
        ska[type][idx] = a.kin.y[gi];
        skc[type][idx] = a.c[gi];
    }
    __syncthreads();
    // interval coefficients (cudaICEEMDAN.cu:394-398)
    for (int i = tid; i < cnt0 + cnt1 - 2; i += kThreads) {
        const int type = (i >= cnt0 - 1);
        const int k = type ? i - (cnt0 - 1) : i;
        const real hk = (real)(skp[type][k + 1] - skp[type][k]);
        const real c0 = skc[type][k], c1 = skc[type][k + 1];
        skb[type][k] = (ska[type][k + 1] - ska[type][k]) / hk - hk * ((real)2 * c0 + c1) / (real)3;
        skd[type][k] = (c1 - c0) / ((real)3 * hk);
    }
    __syncthreads();

    const bool use_mean = (n0 > 3) && (n1 > 3);
    int sxc = 0, sxm = 0;
    for (int v = tid; v < g.nt + 2; v += kThreads) {
        const int p = g.t0 - 1 + v;
        if (p < 0 || p >= g.N) continue;
        const int ku = interval_of(0, p) - klo0;
        const int kl = interval_of(1, p) - klo1;
        const real tu = (real)(p - skp[0][ku]);
        const real U = ska[0][ku] + (skb[0][ku] + (skc[0][ku] + skd[0][ku] * tu) * tu) * tu;
        const real tl = (real)(p - skp[1][kl]);
        const real L = ska[1][kl] + (skb[1][kl] + (skc[1][kl] + skd[1][kl] * tl) * tl) * tl;
        const real m = use_mean ? (U + L) / (real)2 : (real)0;
        const real hn = sh[v + 1] - m;
        sh[v + 1] = hn;
        if (v >= 1 && v <= g.nt) {
            hout[p] = hn;
            if (a.sx_cnt) {
                const real sx = fabs((U + L) / (U - L));
                sxc += (sx > (real)0.05) ? 1 : 0;
                sxm |= (sx > (real)0.5) ? 1 : 0;
            }
        }
    }
    if (a.sx_cnt) {
        // workgroup sums of the Rilling partials (integers: order-free)
        for (int o = 32; o >= 1; o >>= 1) {
            sxc += __shfl_xor(sxc, o, kWave);
            sxm |= __shfl_xor(sxm, o, kWave);
        }
        if ((tid & 63) == 0) {
            s_red[2 * (tid >> 6)] = sxc;
            s_red[2 * (tid >> 6) + 1] = sxm;
        }
        __syncthreads();
        if (tid == 0) {
            int c = 0, mk = 0;
            for (int w = 0; w < kThreads / kWave; ++w) {
                c += s_red[2 * w];
                mk |= s_red[2 * w + 1];
            }
            a.sx_cnt[gt] = c;
            a.sx_mark[gt] = mk;
        }
    }
    if (!a.emit) return;
    __syncthreads();
    emit_tile<real, T>(sd, a.kout, a.pc, g, sh, mask, cpre, &s_misc[4]);
}

// ------------------------------------------------------- windowed spline solve

// Row g of the natural-spline system (cudaICEEMDAN.cu:310-343); rows outside
// [0, n) and the two end rows are identity rows with a zero right-hand side.
// LDS index i holds knot g0 - 1 + i.
template <typename real>
__device__ __forceinline__ void spline_row(int n, long g, int i, const int* sp, const real* sy,
                                           real& dl, real& d, real& du, real& f)
{
    if (g <= 0 || g >= n - 1) {
        dl = (real)0; d = (real)1; du = (real)0; f = (real)0;
        return;
    }
    const int hl = sp[i] - sp[i - 1];
    const int hr = sp[i + 1] - sp[i];
    dl = (real)hl;
    du = (real)hr;
    d = (real)(2 * (hl + hr));
    f = (real)3 * ((sy[i + 1] - sy[i]) / (real)hr - (sy[i] - sy[i - 1]) / (real)hl);
}

template <typename real>
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

template <typename real, int B, int H>
__global__ __launch_bounds__(kThreads) void k_solve(SolveArgs<real> a)
{
    constexpr int W = kWave * B;
    constexpr int K = W - 2 * H;
    constexpr int NW = kThreads / kWave;
    __shared__ int s_pos[NW][W + 2];
    __shared__ real s_y[NW][W + 2];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int* sp = s_pos[wv];
    real* sy = s_y[wv];
    const unsigned total = *a.wincount;
    for (unsigned item = blockIdx.x * NW + wv; item < total; item += gridDim.x * NW) {
        const int s = a.winlist[2 * item];
        const int code = a.winlist[2 * item + 1];
        const int type = code & 1;
        const int w = code >> 1;
        if (a.active && !a.active[s]) continue;
        const int n = a.k.n[2 * s + type];
        const int64_t kb = a.sd.kbase[s] + (type ? a.sd.kstride : 0);
        const long g0 = (long)w * K - H;
        const long lo = g0 - 1;
        for (int i = lane; i < W + 2; i += kWave) {
            const long gg = lo + i;
            if (gg >= 0 && gg < n) {
                sp[i] = a.k.x[kb + gg];
                sy[i] = a.k.y[kb + gg];
            }
        }
        wave_sync<real>();
        // endpoint fix-up when n >= 6 (cudaICEEMDAN.cu:209-291); idempotent,
        // so windows that overlap an end may all apply it
        if (lane == 0 && n >= 6) {
            if (lo <= 0) {
                const int i0 = (int)(-lo);
                const real slope = (sy[i0 + 2] - sy[i0 + 1]) / (real)(sp[i0 + 2] - sp[i0 + 1]);
                const real t = sy[i0 + 1] - slope * (real)(sp[i0 + 1] - sp[i0]);
                if (type == 0 ? (t > sy[i0]) : (t < sy[i0])) sy[i0] = t;
                if (w == 0) a.k.y[kb] = sy[i0];
            }
            const long ie = (long)(n - 1) - lo;
            if (ie >= 0 && ie < W + 2) {
                const int i1 = (int)ie;
                const real slope = (sy[i1 - 1] - sy[i1 - 2]) / (real)(sp[i1 - 1] - sp[i1 - 2]);
                const real t = sy[i1 - 1] + slope * (real)(sp[i1] - sp[i1 - 1]);
                if (type == 0 ? (t > sy[i1]) : (t < sy[i1])) sy[i1] = t;
                const long keep0 = (long)w * K;
                if (n - 1 >= keep0 && n - 1 < keep0 + K) a.k.y[kb + n - 1] = sy[i1];
            }
        }
        wave_sync<real>();

        // per-lane elimination of the interior rows q = 0..B-2:
        // x_q = al_q + be_q * z_{lane-1} + ga_q * z_lane
        real cp[B], al[B], be[B], ga[B];
        const int rb = lane * B;
#pragma unroll
        for (int q = 0; q <= B - 2; ++q) {
            real dl, d, du, f;
            spline_row<real>(n, g0 + rb + q, rb + q + 1, sp, sy, dl, d, du, f);
            real inv;
            if (q == 0) {
                const real dl0 = (lane == 0) ? (real)0 : dl;
                inv = (real)1 / d;
                al[0] = f * inv;
                be[0] = (-dl0) * inv;
            } else {
                inv = (real)1 / (d - dl * cp[q - 1]);
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
        spline_row<real>(n, g0 + rb + B - 1, rb + B, sp, sy, dl, d, du, f);
        real na0 = shfl_down(al[0], 1), nb0 = shfl_down(be[0], 1), ng0 = shfl_down(ga[0], 1);
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
        // parallel cyclic reduction across the wave
#pragma unroll
        for (int st = 1; st < kWave; st <<= 1) {
            const real Am = shfl_up(A, st), Dm = shfl_up(D, st), Cm = shfl_up(C, st),
                       Rm = shfl_up(R, st);
            const real Ap = shfl_down(A, st), Dp = shfl_down(D, st), Cp = shfl_down(C, st),
                       Rp = shfl_down(R, st);
            real a2 = (real)0, c2 = (real)0, d2 = D, r2 = R;
            if (lane - st >= 0) {
                const real k1 = A / Dm;
                a2 = -(Am * k1);
                d2 = d2 - Cm * k1;
                r2 = r2 - Rm * k1;
            }
            if (lane + st < kWave) {
                const real k2 = C / Dp;
                c2 = -(Cp * k2);
                d2 = d2 - Ap * k2;
                r2 = r2 - Rp * k2;
            }
            A = a2;
            C = c2;
            D = d2;
            R = r2;
        }
        const real z = R / D;
        real zl = shfl_up(z, 1);
        if (lane == 0) zl = (real)0;
        real* cout = a.c + kb;
#pragma unroll
        for (int q = 0; q < B; ++q) {
            const int r = rb + q;
            const long gg = g0 + r;
            if (r >= H && r < H + K && gg < n) {
                const real x = (q == B - 1) ? z : (al[q] + be[q] * zl) + ga[q] * z;
                cout[gg] = x;
            }
        }
        wave_sync<real>();
    }
}

// ------------------------------------------------------------ Rilling decide

// siftingCriterion (cudaICEEMDAN.cu:494-525) + checkBreak (:546-580) with
// per-series masks: export at the first stop or at pass cap-1.
template <typename real>
__global__ void k_rilling_decide(SeriesDev sd, const int* nk, const int* sx_cnt,
                                 const int* sx_mark, int* active, int* stop_pass, int* subs,
                                 int pass, int cap, unsigned* n_active)
{
    const int s = blockIdx.x;
    const int lane = threadIdx.x;
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

template <typename real>
__global__ void k_sub(real* w, const real* e, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w[i] = w[i] - e[i];
}

template <typename real>
__global__ void k_copy(real* dst, const real* src, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
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

__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_check_index(const int* x, int64_t N, int* err)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N && x[i] != (int)i) atomicOr(err, ICM_DEVERR_INDEX);
}

// ---------------------------------------------------------------- launchers

static inline unsigned blocks_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

template <typename real> struct Geom;
template <> struct Geom<double> { static constexpr int T = 512, B = 12, H = 64; };
template <> struct Geom<float> { static constexpr int T = 512, B = 8, H = 32; };

template <typename real>
hipError_t launch_init(const SeriesDev& sd, const InitArgs<real>& ia, real* h_out,
                       const KnotsDev<real>& kout, const PassCtl& pc, hipStream_t st)
{
    hipLaunchKernelGGL((k_init<real, Geom<real>::T>), dim3(sd.ntiles), dim3(kThreads), 0, st, sd,
                       ia, h_out, kout, pc);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_sift(const SiftArgs<real>& a, hipStream_t st)
{
    hipLaunchKernelGGL((k_sift<real, Geom<real>::T>), dim3(a.sd.ntiles), dim3(kThreads), 0, st, a);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_solve(const SolveArgs<real>& a, int grid, hipStream_t st)
{
    hipLaunchKernelGGL((k_solve<real, Geom<real>::B, Geom<real>::H>), dim3(grid), dim3(kThreads),
                       0, st, a);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_rilling_decide(const SeriesDev& sd, const int* nk, const int* sx_cnt,
                                 const int* sx_mark, int* active, int* stop_pass, int* subs,
                                 int pass, int cap, unsigned* n_active, hipStream_t st)
{
    hipLaunchKernelGGL((k_rilling_decide<real>), dim3(sd.nseries), dim3(kWave), 0, st, sd, nk,
                       sx_cnt, sx_mark, active, stop_pass, subs, pass, cap, n_active);
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
    hipLaunchKernelGGL((k_sub<real>), dim3(blocks_for(n, 256)), dim3(256), 0, st, w, e, n);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_copy(real* dst, const real* src, int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL((k_copy<real>), dim3(blocks_for(n, 256)), dim3(256), 0, st, dst, src, n);
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

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n, 256)), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_check_index(const int* d_x, int64_t N, int* err, hipStream_t st)
{
    hipLaunchKernelGGL(k_check_index, dim3(blocks_for(N, 256)), dim3(256), 0, st, d_x, N, err);
    return hipGetLastError();
}

#define ICM_INSTANTIATE(real)                                                                      \
    template hipError_t launch_init<real>(const SeriesDev&, const InitArgs<real>&, real*,          \
                                          const KnotsDev<real>&, const PassCtl&, hipStream_t);     \
    template hipError_t launch_sift<real>(const SiftArgs<real>&, hipStream_t);                     \
    template hipError_t launch_solve<real>(const SolveArgs<real>&, int, hipStream_t);              \
    template hipError_t launch_rilling_decide<real>(const SeriesDev&, const int*, const int*,      \
                                                    const int*, int*, int*, int*, int, int,        \
                                                    unsigned*, hipStream_t);                       \
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

ICM_INSTANTIATE(float)
ICM_INSTANTIATE(double)

}  // namespace icm
