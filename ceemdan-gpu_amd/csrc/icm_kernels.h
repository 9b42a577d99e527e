// Kernel argument blocks and launch helpers of the ICEEMDAN HIP path.
//
// Data layout in HBM (one "series" = one noise realization of one channel, or
// one noise realization of a noise group):
//   samples   real[sum_s len_s]            series s at off[s], contiguous
//   knots     int x[2][kstride], real y[2][kstride], real c[2][kstride]
//             type 0 = maxima, 1 = minima; series s owns [kbase[s], kbase[s]+len_s/2+2)
//   tiles     T samples each; series s owns global tiles [tile0[s], tile0[s]+ntl[s])
//   kpre      int[ntiles][2]   knots before the tile (exclusive prefix in the series)
//   nk        int[nseries][2]  knot counts of the series
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace icm {

constexpr int kThreads = 256;   // workgroup size of the tile kernels (4 waves)

struct SeriesDev {
    const int* tile_order;    // [ntiles] launch order -> global tile (tile-major across series)
    const int* tile_series;   // [ntiles]
    const int* tile0;         // [nseries]
    const int* ntl;           // [nseries]
    const int64_t* off;       // [nseries]
    const int* len;           // [nseries]
    const int64_t* kbase;     // [nseries] offset of the series in the compact c array
    const int* chan;          // [nseries] channel (signal phase) or noise group (noise phase)
    const int64_t* eoff;      // [nseries] offset of the series' noise-mode row
    const int* nidx;          // [nseries] noise series whose modes this series uses
    int nseries;
    int ntiles;
    int max_tiles;            // most tiles of one series
    int64_t kstride;          // per-type stride of the compact c array
    int cap;                  // most knots of one type in a tile (T/2 + 2)
};

// Knots of one sift-pass input.  A knot is a set bit of its tile's row
// bitmasks, written by the kernel that wrote the samples (stage start or the
// previous sift pass); its position is the bit's sample index and its value
// the sample of the pass input there.  Nothing else is stored per knot.
template <typename real>
struct KnotsDev {
    uint64_t* mask;   // [ntiles][2][4] knot bits of the tile's four 64-sample rows per type
    int* cnt;    // [ntiles][2]  knots in the tile (popcounts of its masks)
    int* pre;    // [ntiles][2]  knots before the tile in its series (k_knot_prefix)
    int* n;      // [nseries][2] knots of the series (k_knot_prefix)
    real* fix;   // [nseries][2][2] values of knots 0 and n-1 after the endpoint fix (k_solve)
    int* nbx;    // [ntiles][2][3] positions of the knots pre-1, pre-2 and pre+cnt of the
    real* nby;   //  tile (-1 = none) and their values (k_knot_prefix)
};

// A run of consecutive tiles of one series, sifted by one workgroup that
// prefetches tile i+1 while it computes tile i.
struct Segment {
    int s;        // series
    int t_first;  // first tile (within the series)
    int count;    // tiles
    int pad;
};

// window item of the solve: series, (window << 1) | type, first tile or -1
constexpr int kWinItem = 3;

struct PassCtl {
    uint32_t epoch;        // launch id (unique within one forward)
    int* winlist;          // [kWinItem * maxwin] window items for the next solve
    unsigned* wincount;    // [epochs] number of window items
    int K;                 // rows kept per window of the solve
    int H;                 // halo rows on each side of a window
    int* err;              // device error word
    unsigned* dense;       // [epochs] tiles whose knots overflow the small sift table
    // RILLING: series of the phase still sifting after the previous pass
    // (null in FIXED mode); 0 makes the pass's kernels return at once, so a
    // fixed number of passes can be launched (or captured) without asking
    // the host whether any series is left
    const unsigned* run;
};

enum InitMode : int {
    INIT_COPY = 0,     // h0 = src[off + p]                       (noise residual)
    INIT_RESID = 3,    // h0 = src[off + p] - E[off + p], stored in resid[off + p] (the
                       //  next noise residual, formed where it is first read) and in
                       //  h_out unless h_out is null (the first pass then reads resid)
    INIT_STAGE0 = 1,   // h0 = xt[c][p] + (eps0 * E[p]) * inv[s]  (cudaICEEMDAN.cu:68-78)
    INIT_STAGEK = 2,   // h0 = r[c][p] + (eps0 * E[p]) * sq[c]    (cudaICEEMDAN.cu:80-90)
    INIT_KNOTS = 4,    // knots of src only (no samples written: the pass reads src itself)
};

template <typename real>
struct InitArgs {
    int mode;
    const real* src;        // INIT_COPY / INIT_RESID: samples buffer (series offsets)
    real* resid;            // INIT_RESID: h0 also stored here (another buffer than src)
    const real* chan_sig;   // INIT_STAGE0/K: per-channel signal (xt or r)
    const int64_t* chan_off;  // offset of channel c in chan_sig
    const real* E;          // noise modes of this stage (eoff offsets); INIT_RESID: the
                            //  previous noise mode (series offsets)
    const double* scale;    // INIT_STAGE0: inv[s] per series; INIT_STAGEK: sq[c] per channel
    real eps0;
};

template <typename real>
struct SiftArgs {
    SeriesDev sd;
    const Segment* segs;    // [nsegs] one workgroup each
    const real* h_in;
    real* h_out;
    KnotsDev<real> kin;     // knots of h_in
    const real* c;          // spline coefficients c of kin
    KnotsDev<real> kout;    // knots of h_out (written when emit != 0)
    int emit;               // write knots of h_out
    PassCtl pc;
    const int* active;      // Rilling: per-series active flag (nullptr = all)
    int* sx_cnt;            // Rilling: per-tile count of sx > 0.05 (nullptr = off)
    int* sx_mark;           // Rilling: per-tile any(sx > 0.5)
    // sparse pass (non-null): sift with small knot tables, list the tiles
    // that do not fit in ovf_segs[*ovf_count] (zero before the launch) and
    // sift those in a second launch with full tables
    Segment* ovf_segs;      // [ntiles]
    unsigned* ovf_count;
    // small tables only if at most dense_max tiles of the pass input overflow
    // them (*dense, counted by the knot prefix); else full tables throughout
    const unsigned* dense;
    unsigned dense_max;
    // work queue of the pass (zero before the launch; null: one wavefront
    // per segment): a resident grid of wavefronts takes segments from it one
    // at a time until none is left
    unsigned* queue;
};

template <typename real>
struct SolveArgs {
    SeriesDev sd;
    KnotsDev<real> k;       // knots of the pass input
    const real* h;          // the pass input samples (the knots' values)
    real* c;                // [2][kstride] out
    const int* winlist;
    const unsigned* wincount;
    const int* active;
    int* err;
    int* fixf;              // optional [nseries][2][2]: endpoint fix applied (for the adjoint)
    int nak;                // not-a-knot spline ends
    const unsigned* run;    // RILLING: as PassCtl::run
};

// ---------------------------------------------------------------- adjoint
// Recorded structure of one forward pass (the knots of its input samples).
struct PassRec {
    const uint64_t* mask;   // [ntiles][2][4] knot bitmasks
    const int* left;        // [ntiles][2] position of the last knot before the tile (-1: none)
    const int* right;       // [ntiles][2] position of the first knot after the tile (-1: none)
    const int* cnt;         // [ntiles][2]
    const int* pre;         // [ntiles][2]
    const int* n;           // [nseries][2]
    const int* fixf;        // [nseries][2][2]
};

template <typename real>
struct AdjArgs {
    SeriesDev sd;
    PassRec ps;
    real* mu;             // gradient w.r.t. the pass output; becomes the pass input's
    real* mom;            // [2][kstride][4] per-interval moments of mu
    real* edge;           // [ntiles][2][2][4] moments of the runs at tile edges
    real* ga;             // [2][kstride] gradient w.r.t. the knot values
    const int* winlist;
    const unsigned* wincount;
    int* err;
};

// One launch of the fused per-tile adjoint step (256-sample tiles, one
// wavefront each): optional broadcast of dL/dr_{k+1} into mu, optional
// scatter of the knot gradients of pass `sc` (mu <- mu - G^T/2), optional
// per-interval moments of the resulting mu over the knots of pass `mo`.
template <typename real>
struct AdjTileArgs {
    SeriesDev sd;
    real* mu;
    const real* bcast;          // non-null: mu = bcast[chan_off[chan[s]] + p] first
    const int64_t* chan_off;
    PassRec sc;                 // scatter pass (sc.mask == nullptr: none)
    const real* ga;             // knot gradients of pass sc
    PassRec mo;                 // moments pass (mo.mask == nullptr: none)
    real* mom;
    real* edge;
};

// loop = true takes the grid-stride instance even where one wavefront per
// tile fits a launch (for testing that instance)
template <typename real>
hipError_t launch_adj_tile(const AdjTileArgs<real>& a, hipStream_t st, bool loop = false);
template <typename real>
hipError_t launch_adj_span_knot(const AdjArgs<real>& a, int solve_grid, hipStream_t st);
template <typename real>
hipError_t launch_adj_seed(const real* gimf, int M, int64_t N, int64_t stride, real* gr,
                           hipStream_t st);
template <typename real>
hipError_t launch_adj_stage(int stage, int nchan, int I, int max_len, const int* chan_len,
                            const int64_t* chan_off, const SeriesDev& sd, const real* gr_next,
                            const real* mu, const real* E, const real* r_k, real eps0, real* gr_k,
                            real* gacc, real* dots, hipStream_t st);
template <typename real>
hipError_t launch_adj_input(int nchan, const int* chan_len, const int64_t* chan_off, const real* y,
                            const real* gxt, real* gy, hipStream_t st);

// host-side launchers (icm_kernels.hip)
// loop = true: the grid-stride instance even where one wavefront per run of
// tiles fits a launch (for testing that instance)
template <typename real>
hipError_t launch_init(const SeriesDev& sd, const InitArgs<real>& ia, real* h_out,
                       const KnotsDev<real>& kout, const PassCtl& pc, hipStream_t st,
                       bool loop = false);
template <typename real>
hipError_t launch_sift(const SiftArgs<real>& a, int nsegs, hipStream_t st);
template <typename real>
hipError_t launch_solve(const SolveArgs<real>& a, int grid, hipStream_t st);

template <typename real>
hipError_t launch_knot_prefix(const SeriesDev& sd, const KnotsDev<real>& k, const real* h,
                              const PassCtl& pc, const int* active, hipStream_t st);
template <typename real>
hipError_t launch_rilling_decide(const SeriesDev& sd, const int* nk, const int* sx_cnt,
                                 const int* sx_mark, int* active, int* stop_pass, int* subs,
                                 int pass, int cap, const unsigned* run, unsigned* n_active,
                                 hipStream_t st);
template <typename real>
hipError_t launch_series_var(const real* base, const int64_t* off, int64_t stride,
                             const int* len, int nser, real* var_out, hipStream_t st);
template <typename real>
hipError_t launch_noise(uint64_t seed, int I, int64_t N, real* out, int64_t row_stride,
                        hipStream_t st);
template <typename real>
hipError_t launch_ensemble(int stage, int nchan, int I, int max_len, const int* chan_len,
                           const int64_t* chan_off, const int64_t* ser_off,
                           const int64_t* ser_eoff, const int* ser_nidx, const real* hA, const real* hB,
                           const int* mode_parity, int fixed_parity, const real* E,
                           const double* inv, const double* sq, real eps0, const real* xt,
                           real* r, real* imf_base, const int64_t* chan_imf_off, hipStream_t st);
template <typename real>
hipError_t launch_standardize(const real* y, const real* var, real* xt, int64_t N,
                              hipStream_t st);
template <typename real>
hipError_t launch_inv_sqrt(const real* var, double* out, int n, int take_sqrt, hipStream_t st);
template <typename real>
hipError_t launch_sub(real* w, const real* e, int64_t n, hipStream_t st);
template <typename real>
hipError_t launch_copy(real* dst, const real* src, int64_t n, hipStream_t st);
template <typename real>
hipError_t launch_select_copy(const SeriesDev& sd, int max_len, const real* hA, const real* hB,
                              const int* parity, real* dst, hipStream_t st);
hipError_t launch_rec_left(const int* nbx, int* left, int* right, int ntiles, hipStream_t st);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t st);
hipError_t launch_knot_digest(const int* tile_series, const int* tile0, const uint64_t* mask,
                              int ntiles, uint64_t* out, int nseries, hipStream_t st);
hipError_t launch_check_index(const int* d_x, int64_t N, int* err, hipStream_t st);

}  // namespace icm
