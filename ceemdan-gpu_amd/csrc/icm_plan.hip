// Host orchestration of the ICEEMDAN forward (C ABI in include/iceemdan_mi355x.h).
//
// Replaces the reference host drivers iceemdanS/iceemdanD
// (cudaICEEMDAN.cu:619-1338 and :1340-2059).  Differences by design:
//  * one plan owns every buffer; nothing is allocated per call
//  * many channels of uneven length and all their realizations are sifted by
//    the same launches ("series" = one realization of one channel)
//  * FIXED mode has no host round trip at all and is captured once in a
//    hipGraph; RILLING mode masks finished series on the device and polls a
//    single counter for the early exit (the reference syncs 5x per pass)
//  * noise modes are computed once per distinct length and shared by every
//    channel of that length (all channels use the same seed, as every
//    reference call uses seed 1234, cudaICEEMDAN.cu:740)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/iceemdan_mi355x.h"
#include "icm_errors.h"
#include "icm_kernels.h"

namespace icm {

// text of the last HIP failure seen by an entry point on this thread
thread_local char g_last_error[256] = "";

int hip_fail(hipError_t e, int line)
{
    std::snprintf(g_last_error, sizeof g_last_error, "icm_plan.hip:%d: %s", line,
                  hipGetErrorString(e));
    return ICM_ERR_HIP;
}

const int64_t kMaxLen = 16000000;
const uint32_t kEpochs = 65536;
// Passes per stage: one epoch per pass (FIXED uses S + 1 epochs per stage,
// RILLING at most cap + 1, and the control words are reset when fewer than
// 2 cap + 8 remain), so a cap must leave a stage's epochs below kEpochs.
const int kMaxSifts = 32000;
// Channels: the per-channel launches put the channel on gridDim.y.
const int kMaxChannels = 65535;
// tiles of one phase (series x tiles per series)
const int64_t kMaxTiles = (int64_t)1 << 28;

// --------------------------------------------------------------- geometry

// samples per tile of the tile kernels (one wavefront per tile)
int tile_of(int) { return 256; }
// windowed solve geometry: 64 lanes x 8 rows and H halo rows on each side.
// A row couples to its neighbours by hl and hr against 2 (hl + hr), so what
// a window drops decays by 3.7x per row at uniform spacing (slower only
// along runs of spacings that shrink geometrically).  FP64 keeps 32 rows
// (its windows equal a global solve to round-off); FP32 keeps 16: against a
// global solve 6e-11 of the largest coefficient on EEG extrema and 7e-11 at
// uniform spacing, far below FP32 rounding (the worst-case bound, 2^(1-H)
// from the Jacobi matrix's infinity norm of 1/2, is 3e-5 at H = 16; only
// runs of geometrically shrinking spacings approach it).
int solve_B(int) { return 8; }
int solve_H(int dtype) { return dtype == ICM_DTYPE_F32 ? 16 : 32; }
int solve_K(int dtype) { return 64 * solve_B(dtype) - 2 * solve_H(dtype); }

int validate(const icm_config* c)
{
    if (!c) return ICM_ERR_INVALID_ARG;
    if (c->dtype != ICM_DTYPE_F32 && c->dtype != ICM_DTYPE_F64) return ICM_ERR_INVALID_ARG;
    if (c->num_channels < 1 || c->num_channels > kMaxChannels || !c->lengths)
        return ICM_ERR_INVALID_ARG;
    for (int i = 0; i < c->num_channels; ++i)
        if (c->lengths[i] < 4 || c->lengths[i] > kMaxLen) return ICM_ERR_INVALID_ARG;
    if (c->num_noise < 1 || c->num_noise > 65535) return ICM_ERR_INVALID_ARG;
    if (c->num_imfs < 2 || c->num_imfs > 64) return ICM_ERR_INVALID_ARG;
    if (c->stop_mode != ICM_STOP_FIXED && c->stop_mode != ICM_STOP_RILLING) return ICM_ERR_INVALID_ARG;
    if (c->max_sifts < 1 || c->max_sifts > kMaxSifts) return ICM_ERR_INVALID_ARG;
    if (c->noise_max_sifts < 1 || c->noise_max_sifts > kMaxSifts) return ICM_ERR_INVALID_ARG;
    if (!std::isfinite(c->noise_strength)) return ICM_ERR_INVALID_ARG;
    if (c->boundary != ICM_BOUNDARY_NATURAL && c->boundary != ICM_BOUNDARY_NOT_A_KNOT)
        return ICM_ERR_INVALID_ARG;
    int64_t total = 0;
    for (int i = 0; i < c->num_channels; ++i) total += c->lengths[i];
    if (total * (int64_t)c->num_noise > (int64_t)1 << 40) return ICM_ERR_INVALID_ARG;
    // tiles of a phase are indexed with 32-bit integers (eight mask words per
    // tile and type pairs included): at most 2^28 tiles.  Launch grids are
    // capped (grid-stride and work-queue kernels), so this is the only limit
    // on the work of one launch.
    int64_t tiles = 0;
    const int T = tile_of(c->dtype);
    for (int i = 0; i < c->num_channels; ++i) tiles += (c->lengths[i] + T - 1) / T;
    if (tiles * c->num_noise > kMaxTiles) return ICM_ERR_INVALID_ARG;
    return ICM_OK;
}

// ---------------------------------------------------------- sift geometry
// The sift pass is scheduled as segments (runs of consecutive tiles of one
// series) taken from a work queue.  Segment list 0 is the default geometry
// (clamp(tiles / 56000, 1, 32) tiles, or ICM_SEGMENT_TILES); lists 1..kNSeg
// have kSegCand[l - 1] tiles per segment and are the autotuner's candidates.
// Tuning candidate c (0 .. kNCand-1): list 1 + c / 2, small knot tables if c
// is odd; -1 = the default (list 0, small tables from stage sparse_stage on).
constexpr int kNSeg = 4;
constexpr int kSegCand[kNSeg] = {4, 8, 16, 32};
constexpr int kNCand = 2 * kNSeg;
constexpr int kNLists = 1 + kNSeg;

// Series tables of one phase (host side), built from (length, count) groups.
struct HostSeries {
    std::vector<int> tile_order, tile_series, tile0, ntl, len, chan, nidx;
    std::vector<Segment> segs[kNLists];
    std::vector<int64_t> off, kbase, eoff;
    int64_t samples = 0, kcap = 0, maxwin = 0;
    int ntiles = 0;
    int max_len = 0;
};

static void build_series(HostSeries& hs, const std::vector<int64_t>& lens, int I, int T, int K,
                         const std::vector<int64_t>* noise_off,   // per channel: noise row base
                         const std::vector<int>* noise_grp)
{
    int64_t off = 0, kb = 0;
    int tiles = 0;
    for (size_t c = 0; c < lens.size(); ++c) {
        const int L = (int)lens[c];
        hs.max_len = std::max(hs.max_len, L);
        for (int i = 0; i < I; ++i) {
            const int s = (int)hs.len.size();
            const int nt = (L + T - 1) / T;
            hs.len.push_back(L);
            hs.off.push_back(off);
            hs.kbase.push_back(kb);
            hs.tile0.push_back(tiles);
            hs.ntl.push_back(nt);
            hs.chan.push_back((int)c);
            if (noise_off) {
                const int g = (*noise_grp)[c];
                hs.eoff.push_back((*noise_off)[g] + (int64_t)i * L);
                hs.nidx.push_back(g * I + i);
            } else {
                hs.eoff.push_back(off);
                hs.nidx.push_back(s);
            }
            for (int t = 0; t < nt; ++t) hs.tile_series.push_back(s);
            tiles += nt;
            off += L;
            const int64_t cap = L / 2 + 2;
            kb += cap;
            hs.maxwin += 2 * ((cap + K - 1) / K);
        }
    }
    hs.samples = off;
    hs.kcap = kb;
    hs.ntiles = tiles;
    // segments of consecutive tiles (one wavefront each for 256-sample
    // tiles, one workgroup each otherwise): long enough to amortize the
    // prefetch pipeline, short enough to fill every CU
    int seg0 = 1;
    const char* env = std::getenv("ICM_SEGMENT_TILES");
    if (env && std::atoi(env) > 0) seg0 = std::atoi(env);
    else seg0 = std::max(1, std::min(32, tiles / (T == 256 ? 56000 : 6144)));
    for (int l = 0; l < kNLists; ++l) {
        const int seg = l == 0 ? seg0 : kSegCand[l - 1];
        for (size_t s = 0; s < hs.ntl.size(); ++s)
            for (int t = 0; t < hs.ntl[s]; t += seg) {
                Segment g;
                g.s = (int)s;
                g.t_first = t;
                g.count = std::min(seg, hs.ntl[s] - t);
                g.pad = 0;
                hs.segs[l].push_back(g);
            }
    }
    // launch order of the tile kernels: tile t of every series before tile t+1
    int maxt = 0;
    for (int nt : hs.ntl) maxt = std::max(maxt, nt);
    hs.tile_order.reserve(tiles);
    for (int t = 0; t < maxt; ++t)
        for (size_t s = 0; s < hs.ntl.size(); ++s)
            if (t < hs.ntl[s]) hs.tile_order.push_back(hs.tile0[s] + t);
}

// --------------------------------------------------------- tuning table
// Chosen sift candidates per stage, by plan shape, shared by the plans of
// this process (a plan of a shape already tuned starts tuned).
std::mutex g_tune_mu;
std::map<std::string, std::vector<int>> g_tune_table;

bool tune_table_get(const std::string& key, std::vector<int>& out)
{
    std::lock_guard<std::mutex> lock(g_tune_mu);
    auto it = g_tune_table.find(key);
    if (it == g_tune_table.end() || it->second.size() != out.size()) return false;
    out = it->second;
    return true;
}

void tune_table_put(const std::string& key, const std::vector<int>& v)
{
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune_table[key] = v;
}

// ------------------------------------------------------------------ plan

struct DevSeries {
    Segment* segs[kNLists] = {};
    int nsegs[kNLists] = {};
    int *tile_order = nullptr, *tile_series = nullptr, *tile0 = nullptr, *ntl = nullptr, *len = nullptr, *chan = nullptr,
        *nidx = nullptr;
    int64_t *off = nullptr, *kbase = nullptr, *eoff = nullptr;
    int nseries = 0, ntiles = 0, max_len = 0, max_tiles = 0;
    int64_t samples = 0, kcap = 0, maxwin = 0;
    SeriesDev view(int64_t kstride, int cap) const
    {
        SeriesDev v{};
        v.cap = cap;
        v.tile_order = tile_order; v.tile_series = tile_series; v.tile0 = tile0; v.ntl = ntl; v.off = off; v.len = len;
        v.kbase = kbase; v.chan = chan; v.eoff = eoff; v.nidx = nidx;
        v.nseries = nseries; v.ntiles = ntiles; v.kstride = kstride;
        v.max_tiles = max_tiles;
        return v;
    }
};

struct PlanBase {
    icm_config cfg{};
    std::vector<int64_t> lens;
    std::vector<int64_t> glen;         // noise-group lengths
    std::vector<int> chan_group;       // channel -> noise group
    std::vector<int64_t> group_off;    // group -> first noise sample (I rows)
    std::vector<int64_t> chan_off;     // channel -> offset in per-channel arrays
    std::vector<int64_t> chan_imf_off; // channel -> offset of [M][N] output
    HostSeries hn, hsig;
    DevSeries dn, dsig;
    int64_t kstride = 0, maxwin = 0, ntiles_max = 0, nser_max = 0, samples_max = 0;
    int cap = 0;           // most knots of one type in a tile (T/2 + 2)
    int64_t chan_total = 0;
    std::vector<void*> allocs;
    int64_t bytes = 0;
    hipStream_t own = nullptr;
    // control
    unsigned* wincount = nullptr;
    int* winlist = nullptr;
    int* err = nullptr;
    unsigned* n_active = nullptr;     // [2] RILLING: series still sifting, by pass parity
    unsigned* h_n_active = nullptr;   // [2] pinned copies read by the host (one per poll slot)
    hipEvent_t poll_ev[2] = {nullptr, nullptr};
    bool capturing = false;           // a phase is being captured into a graph
    // the two phases side by side: the noise-mode phase's stream, the
    // events after each of its stages, the fork and the CEEMDAN phase's start
    hipStream_t aux = nullptr;
    std::vector<hipEvent_t> stage_ev;
    hipEvent_t ev_fork = nullptr, ev_sig = nullptr;
    bool overlap_on = true;
    bool sift_queue = false;          // every sift pass from a work queue (ICM_SIFT_QUEUE=1)
    bool grid_loop = false;           // grid-stride start / adjoint tile (ICM_GRID_LOOP=1)
    int* h_err = nullptr;             // pinned
    // rilling
    int *active = nullptr, *stop_pass = nullptr, *subs = nullptr, *sx_cnt = nullptr,
        *sx_mark = nullptr;
    // trace
    int trace_passes = 0;
    int32_t* trace[2] = {nullptr, nullptr};
    uint64_t* trace_hash[2] = {nullptr, nullptr};   // [M-1][passes][series] knot-set digests
    int32_t* stops[2] = {nullptr, nullptr};
    // timing
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    icm_timing timing{};
    // graphs: [0] noise-mode phase, [1] CEEMDAN phase, [2] both side by side
    hipGraphExec_t gexec[3] = {nullptr, nullptr, nullptr};
    bool graph_failed = false;
    uint32_t epoch = 0;
    int64_t* d_chan_off = nullptr;
    int64_t* d_chan_imf_off = nullptr;
    int* d_chan_len = nullptr;
    int max_chan_len = 0;
    bool noise_external = false;
    // per-kernel timing (HIP events around each init / solve / sift launch)
    int prof = 0;
    std::vector<hipEvent_t> prof_ev;
    std::vector<int> prof_kind;
    std::vector<int64_t> prof_bytes;
    size_t prof_used = 0;
    double kern_ms[3] = {0, 0, 0};
    int64_t kern_n[3] = {0, 0, 0};
    int64_t kern_bytes[3] = {0, 0, 0};
    // autotuner of the sift pass: per phase and stage, the segment list and
    // knot-table size (candidate index, -1 = default geometry)
    int tune_left = 0;              // icm_forward calls that still time candidates
    int tune_round = 0;             // tuning forwards completed
    bool tuning = false;            // the running forward times candidates
    std::vector<int> choice[2];     // [M-1]
    std::vector<double> tsum[2];    // [(M-1) * kNCand] ms of timed launches
    std::vector<int> tcnt[2];
    std::vector<hipEvent_t> tev;    // event pairs around timed sift launches
    std::vector<int> ttag;          // (phase, stage, candidate) of each pair
    size_t tused = 0;
    std::string tkey[2];            // shape keys of the process-wide table

    int tune_candidate(int phase, int stage, int pass) const
    {
        if (tuning) return (pass + 3 * tune_round + stage) % kNCand;
        if (stage >= 0 && stage < (int)choice[phase].size()) return choice[phase][stage];
        return -1;
    }
    int tune_begin(int phase, int stage, int cand, hipStream_t st)
    {
        if (!tuning || cand < 0) return ICM_OK;
        while (tev.size() < 2 * (tused + 1)) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return ICM_ERR_HIP;
            tev.push_back(e);
        }
        ttag.resize(tused + 1);
        ttag[tused] = (phase * 65536 + stage) * kNCand + cand;
        return hipEventRecord(tev[2 * tused], st) == hipSuccess ? ICM_OK : ICM_ERR_HIP;
    }
    int tune_end(int cand, hipStream_t st)
    {
        if (!tuning || cand < 0) return ICM_OK;
        const hipError_t e = hipEventRecord(tev[2 * tused + 1], st);
        ++tused;
        return e == hipSuccess ? ICM_OK : ICM_ERR_HIP;
    }
    // after a completed tuning forward (its stream synchronised)
    void tune_collect()
    {
        for (size_t i = 0; i < tused; ++i) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, tev[2 * i], tev[2 * i + 1]) != hipSuccess) continue;
            const int tag = ttag[i];
            const int cand = tag % kNCand, ps = tag / kNCand;
            const int phase = ps / 65536, stage = ps % 65536;
            const size_t at = (size_t)stage * kNCand + cand;
            if (phase < 2 && at < tsum[phase].size()) {
                tsum[phase][at] += ms;
                tcnt[phase][at] += 1;
            }
        }
        tused = 0;
        tuning = false;
        ++tune_round;
        if (tune_left > 0 && --tune_left == 0) tune_finish();
    }
    // per stage, the candidate with the least mean time; stored in the
    // process-wide table under the plan's shape
    void tune_finish()
    {
        tune_left = 0;
        if (tune_round == 0) return;
        for (int ph = 0; ph < 2; ++ph) {
            for (size_t k = 0; k < choice[ph].size(); ++k) {
                int best = -1;
                double bm = 0;
                for (int c = 0; c < kNCand; ++c) {
                    const size_t at = k * kNCand + c;
                    if (!tcnt[ph][at]) continue;
                    const double m = tsum[ph][at] / tcnt[ph][at];
                    if (best < 0 || m < bm) { best = c; bm = m; }
                }
                choice[ph][k] = best;
            }
            tune_table_put(tkey[ph], choice[ph]);
        }
    }
    void drop_graphs()
    {
        for (int i = 0; i < 3; ++i)
            if (gexec[i]) {
                (void)hipGraphExecDestroy(gexec[i]);
                gexec[i] = nullptr;
            }
    }
    // the plan's shape, as the key of the process-wide tuning table
    std::string shape_key(int phase) const
    {
        const HostSeries& h = phase == ICM_PHASE_NOISE ? hn : hsig;
        int dev = 0;
        (void)hipGetDevice(&dev);
        char buf[256];
        std::snprintf(buf, sizeof buf, "d%d|t%d|m%d|b%d|S%d|M%d|p%d|n%zu|T%d|L%d|X%lld", dev,
                      cfg.dtype, cfg.stop_mode, cfg.boundary,
                      phase == ICM_PHASE_NOISE ? cfg.noise_max_sifts : cfg.max_sifts,
                      cfg.num_imfs, phase, h.len.size(), h.ntiles, h.max_len,
                      (long long)h.samples);
        return buf;
    }
    // starts tuned when the table holds this shape; else tunes over the
    // first ICM_AUTOTUNE (default 2) forwards
    void tune_init()
    {
        const int M = cfg.num_imfs;
        for (int ph = 0; ph < 2; ++ph) {
            choice[ph].assign(M - 1, -1);
            tsum[ph].assign((size_t)(M - 1) * kNCand, 0.0);
            tcnt[ph].assign((size_t)(M - 1) * kNCand, 0);
            tkey[ph] = shape_key(ph);
        }
        int n = 2;
        if (const char* env = std::getenv("ICM_AUTOTUNE")) n = std::max(0, std::atoi(env));
        tune_left = n;
        if (n > 0 && tune_table_get(tkey[0], choice[0]) && tune_table_get(tkey[1], choice[1]))
            tune_left = 0;
        else if (tune_left > 0)
            for (int ph = 0; ph < 2; ++ph) choice[ph].assign(M - 1, -1);
    }
    int set_autotune(int forwards)
    {
        if (forwards < 0) return ICM_ERR_INVALID_ARG;
        for (int ph = 0; ph < 2; ++ph) {
            std::fill(tsum[ph].begin(), tsum[ph].end(), 0.0);
            std::fill(tcnt[ph].begin(), tcnt[ph].end(), 0);
            if (forwards == 0) std::fill(choice[ph].begin(), choice[ph].end(), -1);
        }
        tune_left = forwards;
        tune_round = 0;
        drop_graphs();
        return ICM_OK;
    }
    int get_tuning(int phase, int32_t* out, int64_t count) const
    {
        if (phase < 0 || phase > 1 || !out || count < (int64_t)choice[phase].size())
            return ICM_ERR_INVALID_ARG;
        for (size_t k = 0; k < choice[phase].size(); ++k) out[k] = choice[phase][k];
        return ICM_OK;
    }
    int set_tuning(int phase, const int32_t* in, int64_t count)
    {
        if (phase < 0 || phase > 1 || !in || count != (int64_t)choice[phase].size())
            return ICM_ERR_INVALID_ARG;
        for (int64_t k = 0; k < count; ++k)
            if (in[k] < -1 || in[k] >= kNCand) return ICM_ERR_INVALID_ARG;
        for (int64_t k = 0; k < count; ++k) choice[phase][k] = in[k];
        tune_left = 0;
        drop_graphs();
        return ICM_OK;
    }

    int prof_begin(int kind, int64_t bytes, hipStream_t st)
    {
        if (!prof) return ICM_OK;
        while (prof_ev.size() < 2 * (prof_used + 1)) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return ICM_ERR_HIP;
            prof_ev.push_back(e);
        }
        prof_kind.resize(prof_used + 1);
        prof_bytes.resize(prof_used + 1);
        prof_kind[prof_used] = kind;
        prof_bytes[prof_used] = bytes;
        return hipEventRecord(prof_ev[2 * prof_used], st) == hipSuccess ? ICM_OK : ICM_ERR_HIP;
    }
    int prof_end(hipStream_t st)
    {
        if (!prof) return ICM_OK;
        const hipError_t e = hipEventRecord(prof_ev[2 * prof_used + 1], st);
        ++prof_used;
        return e == hipSuccess ? ICM_OK : ICM_ERR_HIP;
    }
    void prof_collect()
    {
        for (int k = 0; k < 3; ++k) { kern_ms[k] = 0; kern_n[k] = 0; kern_bytes[k] = 0; }
        for (size_t i = 0; i < prof_used; ++i) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, prof_ev[2 * i], prof_ev[2 * i + 1]);
            kern_ms[prof_kind[i]] += ms;
            kern_n[prof_kind[i]] += 1;
            kern_bytes[prof_kind[i]] += prof_bytes[i];
        }
        prof_used = 0;
    }

    virtual ~PlanBase() {}
    virtual int forward(const void* const* sig, void* const* imfs, const void* noise,
                        hipStream_t st) = 0;
    virtual int noise_modes_out(void* d_out, int64_t count, hipStream_t st) = 0;
    virtual int backward(const void* const* gimf, void* const* gsig, hipStream_t st) = 0;
    virtual int forward_begin(const void* const* sig, const void* noise, hipStream_t st) = 0;
    virtual int forward_stages(int count, int* stage_out, hipStream_t st) = 0;
    virtual int forward_finish(void* const* out, hipStream_t st) = 0;
    virtual int64_t state_bytes() const = 0;
    virtual int state_save(void* host, int64_t bytes, hipStream_t st) = 0;
    virtual int state_load(const void* host, int64_t bytes, const void* noise, hipStream_t st) = 0;

    template <typename T>
    int alloc(T** p, int64_t n)
    {
        *p = nullptr;
        if (n <= 0) n = 1;
        void* q = nullptr;
        if (hipMalloc(&q, (size_t)n * sizeof(T)) != hipSuccess) return ICM_ERR_OUT_OF_MEMORY;
        allocs.push_back(q);
        bytes += n * (int64_t)sizeof(T);
        *p = (T*)q;
        return ICM_OK;
    }
    template <typename T>
    int upload(T** p, const std::vector<T>& v)
    {
        int rc = alloc(p, (int64_t)v.size());
        if (rc) return rc;
        if (!v.empty() &&
            hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
            return ICM_ERR_HIP;
        return ICM_OK;
    }
    int upload_series(DevSeries& d, const HostSeries& h)
    {
        int rc;
        for (int l = 0; l < kNLists; ++l) {
            if ((rc = upload(&d.segs[l], h.segs[l]))) return rc;
            d.nsegs[l] = (int)h.segs[l].size();
        }
        if ((rc = upload(&d.tile_order, h.tile_order))) return rc;
        if ((rc = upload(&d.tile_series, h.tile_series))) return rc;
        if ((rc = upload(&d.tile0, h.tile0))) return rc;
        if ((rc = upload(&d.ntl, h.ntl))) return rc;
        if ((rc = upload(&d.len, h.len))) return rc;
        if ((rc = upload(&d.chan, h.chan))) return rc;
        if ((rc = upload(&d.nidx, h.nidx))) return rc;
        if ((rc = upload(&d.off, h.off))) return rc;
        if ((rc = upload(&d.kbase, h.kbase))) return rc;
        if ((rc = upload(&d.eoff, h.eoff))) return rc;
        d.nseries = (int)h.len.size();
        d.ntiles = h.ntiles;
        d.max_tiles = 0;
        for (int nt : h.ntl) d.max_tiles = std::max(d.max_tiles, nt);
        d.max_len = h.max_len;
        d.samples = h.samples;
        d.kcap = h.kcap;
        d.maxwin = h.maxwin;
        return ICM_OK;
    }
    void free_all()
    {
        for (int i = 0; i < 3; ++i)
            if (gexec[i]) (void)hipGraphExecDestroy(gexec[i]);
        for (auto& e : stage_ev)
            if (e) (void)hipEventDestroy(e);
        if (ev_sig) (void)hipEventDestroy(ev_sig);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (aux) (void)hipStreamDestroy(aux);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : poll_ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : prof_ev) (void)hipEventDestroy(e);
        for (auto& e : tev) (void)hipEventDestroy(e);
        for (void* p : allocs) (void)hipFree(p);
        allocs.clear();
        if (h_n_active) (void)hipHostFree(h_n_active);
        if (h_err) (void)hipHostFree(h_err);
        if (own) (void)hipStreamDestroy(own);
    }
};

void layout(PlanBase& P)
{
    const icm_config& c = P.cfg;
    const int T = tile_of(c.dtype), K = solve_K(c.dtype), I = c.num_noise;
    // noise groups = distinct lengths, in first-appearance order
    std::map<int64_t, int> gid;
    for (int ch = 0; ch < c.num_channels; ++ch) {
        const int64_t L = c.lengths[ch];
        P.lens.push_back(L);
        auto it = gid.find(L);
        if (it == gid.end()) {
            const int g = (int)P.glen.size();
            gid[L] = g;
            P.glen.push_back(L);
            P.chan_group.push_back(g);
        } else {
            P.chan_group.push_back(it->second);
        }
    }
    int64_t o = 0;
    for (int64_t L : P.glen) {
        P.group_off.push_back(o);
        o += L * I;
    }
    int64_t co = 0, io = 0;
    for (int64_t L : P.lens) {
        P.chan_off.push_back(co);
        P.chan_imf_off.push_back(io);
        co += L;
        io += L * c.num_imfs;
        P.max_chan_len = std::max(P.max_chan_len, (int)L);
    }
    P.chan_total = co;
    build_series(P.hn, P.glen, I, T, K, nullptr, nullptr);
    build_series(P.hsig, P.lens, I, T, K, &P.group_off, &P.chan_group);
    P.kstride = std::max(P.hn.kcap, P.hsig.kcap);
    P.maxwin = std::max(P.hn.maxwin, P.hsig.maxwin);
    P.ntiles_max = std::max(P.hn.ntiles, P.hsig.ntiles);
    P.nser_max = std::max((int64_t)P.hn.len.size(), (int64_t)P.hsig.len.size());
    P.samples_max = std::max(P.hn.samples, P.hsig.samples);
    P.cap = T / 2 + 2;   // strict extrema are never adjacent, plus the two ends
}

int64_t plan_bytes_estimate(const icm_config* c)
{
    PlanBase* dummy = nullptr;
    (void)dummy;
    struct Tmp : PlanBase {
        int forward(const void* const*, void* const*, const void*, hipStream_t) override { return 0; }
        int noise_modes_out(void*, int64_t, hipStream_t) override { return 0; }
        int backward(const void* const*, void* const*, hipStream_t) override { return 0; }
        int forward_begin(const void* const*, const void*, hipStream_t) override { return 0; }
        int forward_stages(int, int*, hipStream_t) override { return 0; }
        int forward_finish(void* const*, hipStream_t) override { return 0; }
        int64_t state_bytes() const override { return 0; }
        int state_save(void*, int64_t, hipStream_t) override { return 0; }
        int state_load(const void*, int64_t, const void*, hipStream_t) override { return 0; }
    } P;
    P.cfg = *c;
    layout(P);
    const int64_t s = c->dtype == ICM_DTYPE_F64 ? 8 : 4;
    const int M = c->num_imfs;
    int64_t b = 0;
    for (const HostSeries* h : {&P.hn, &P.hsig}) {  // pass state of each phase's lane
        const int64_t nt = h->ntiles, ns = (int64_t)h->len.size();
        b += 2 * h->samples * s;                    // h ping-pong
        b += 2 * h->kcap * s;                       // c
        b += 2 * (2 * 2 * nt * 4 + 2 * ns * 4 + 4 * ns * s);   // counts, prefixes, n, fix
        b += 2 * nt * (6 * (4 + s) + 8 * 8);        // neighbour knots, knot bitmasks
        b += (int64_t)kWinItem * h->maxwin * 4;     // window list
        b += 4 * (int64_t)kEpochs * 4;              // control words
        b += nt * (int64_t)sizeof(Segment);         // sparse-pass tile list
    }
    b += (int64_t)(M - 1) * P.hn.samples * s;      // noise modes
    b += 2 * P.hn.samples * s;                     // noise residuals (two buffers)
    for (int l = 0; l < kNLists; ++l)                                         // segment lists
        b += (int64_t)(P.hn.segs[l].size() + P.hsig.segs[l].size()) * (int64_t)sizeof(Segment);
    b += 5 * P.chan_total * s + P.chan_total * M * s;
    b += (int64_t)(M - 1) * P.chan_total * s;     // residue history
    if (c->keep_records && c->stop_mode == ICM_STOP_FIXED)   // per-pass knot records
        b += (int64_t)(M - 1) * c->max_sifts *
             ((int64_t)P.hsig.ntiles * (64 + 16 + 16) + 6 * (int64_t)P.hsig.len.size() * 4 +
              (int64_t)kWinItem * P.hsig.maxwin * 4);
    return b;
}

// ------------------------------------------------------------ typed plan

template <typename real>
struct Plan : PlanBase {
    // Pass state of one phase: the samples' ping-pong buffers, the knots of
    // a pass's input and output, the spline coefficients c, the solve's
    // window list and the per-pass control words (indexed by the lane's
    // epoch, cleared by fill kernels).  The noise-mode phase and the CEEMDAN
    // phase have one each, so that they can run side by side.
    struct Lane {
        real *hA = nullptr, *hB = nullptr;
        int* kcnt[2] = {};
        int* kpre[2] = {};
        int* nk[2] = {};
        real* kfix[2] = {};
        int* knbx[2] = {};
        real* knby[2] = {};
        uint64_t* kmask[2] = {};
        real* kc = nullptr;                // [2][kstride]
        int64_t kstride = 0;               // per-type stride of kc (the phase's knot capacity)
        int* winlist = nullptr;            // [kWinItem * maxwin] window items of the next solve
        unsigned* wincount = nullptr;      // [kEpochs] their number per pass
        unsigned* ovfcount = nullptr;      // [kEpochs] tiles a sparse pass left to the full tables
        unsigned* densecnt = nullptr;      // [kEpochs] tiles of a pass input over the small tables
        unsigned* squeue = nullptr;        // [kEpochs] work-queue counter of each sift launch
        Segment* ovfsegs = nullptr;        // [ntiles] the tiles a sparse pass left
        uint32_t epoch = 0;
    };
    Lane ln[2];                            // [ICM_PHASE_NOISE], [ICM_PHASE_SIGNAL]
    real *modes = nullptr, *wres = nullptr, *wres2 = nullptr;
    real *y = nullptr, *xt = nullptr, *r = nullptr, *imf = nullptr;
    real *var_chan = nullptr, *var_noise = nullptr;
    double *inv_noise = nullptr, *sq_chan = nullptr;
    real* r_hist = nullptr;
    int stage_ = -1;            // next CEEMDAN stage of a staged forward (-1: none in progress)     // [M-1][chan_total]: rows 1 .. M-2 = r_1 .. r_{M-2} of the last forward
    bool have_forward = false;
    // adjoint state (allocated by the first backward)
    struct Rec {
        uint64_t* mask = nullptr;   // [ntiles][2][4] knot bitmasks of the pass input
        int *left = nullptr, *right = nullptr, *cnt = nullptr, *pre = nullptr, *n = nullptr,
            *fixf = nullptr;
        int* winlist = nullptr;
        unsigned* wincount = nullptr;
    };
    int sparse_stage = 1;              // first stage of a phase sifted with small knot tables
    int64_t sparse_dense_max = -1;     // over-full tiles tolerated by them (-1: tiles / 64)
    std::vector<Rec> rec;                  // re-sift records of one stage (backward)
    std::vector<std::vector<Rec>> krec;    // per stage, kept by the forward (keep_records)
    std::vector<char> krec_valid;
    real *mu = nullptr, *mom = nullptr, *edge = nullptr, *ga = nullptr, *gr = nullptr,
         *gacc = nullptr, *dots = nullptr, *gy = nullptr;

    int setup()
    {
        int rc;
        const int M = cfg.num_imfs;
        if ((rc = upload_series(dn, hn))) return rc;
        if ((rc = upload_series(dsig, hsig))) return rc;
        if ((rc = alloc_lane(ln[ICM_PHASE_NOISE], hn))) return rc;
        if ((rc = alloc_lane(ln[ICM_PHASE_SIGNAL], hsig))) return rc;
        if ((rc = alloc(&modes, (int64_t)(M - 1) * hn.samples))) return rc;
        if ((rc = alloc(&wres, hn.samples))) return rc;
        if ((rc = alloc(&wres2, hn.samples))) return rc;
        // ICM_SPARSE_STAGE: first stage sifted with small knot tables (a
        // tuning knob; 0 also sends the dense first stage through them, and
        // most of its tiles through the full-table launch)
        if (const char* env = std::getenv("ICM_SPARSE_STAGE"))
            if (std::atoi(env) >= 0) sparse_stage = std::atoi(env);
        // ICM_SPARSE_DENSE_MAX: over-full tiles a small-table pass tolerates
        // before it runs with full tables throughout (default tiles / 64)
        if (const char* env = std::getenv("ICM_SPARSE_DENSE_MAX"))
            if (std::atoll(env) >= 0) sparse_dense_max = std::atoll(env);
        if ((rc = alloc(&err, 1))) return rc;
        if ((rc = alloc(&n_active, 2))) return rc;
        if ((rc = alloc(&y, chan_total))) return rc;
        if ((rc = alloc(&xt, chan_total))) return rc;
        if ((rc = alloc(&r, chan_total))) return rc;
        if ((rc = alloc(&imf, chan_total * M))) return rc;
        if ((rc = alloc(&r_hist, chan_total * (M - 1)))) return rc;
        if ((rc = alloc(&var_chan, cfg.num_channels))) return rc;
        if ((rc = alloc(&sq_chan, cfg.num_channels))) return rc;
        if ((rc = alloc(&var_noise, (int64_t)hn.len.size()))) return rc;
        if ((rc = alloc(&inv_noise, (int64_t)hn.len.size()))) return rc;
        if ((rc = upload(&d_chan_off, chan_off))) return rc;
        if ((rc = upload(&d_chan_imf_off, chan_imf_off))) return rc;
        std::vector<int> cl(lens.begin(), lens.end());
        if ((rc = upload(&d_chan_len, cl))) return rc;
        if (cfg.stop_mode == ICM_STOP_RILLING) {
            if ((rc = alloc(&active, nser_max))) return rc;
            if ((rc = alloc(&stop_pass, nser_max))) return rc;
            if ((rc = alloc(&subs, nser_max))) return rc;
            if ((rc = alloc(&sx_cnt, ntiles_max))) return rc;
            if ((rc = alloc(&sx_mark, ntiles_max))) return rc;
        }
        if (cfg.keep_records && cfg.stop_mode == ICM_STOP_FIXED) {
            krec.resize(M - 1);
            for (int k = 0; k < M - 1; ++k)
                if ((rc = alloc_recs(krec[k]))) return rc;
            krec_valid.assign(M - 1, 0);
        }
        if (hipHostMalloc((void**)&h_n_active, 2 * sizeof(unsigned), hipHostMallocDefault) !=
            hipSuccess)
            return ICM_ERR_OUT_OF_MEMORY;
        if (hipHostMalloc((void**)&h_err, sizeof(int), hipHostMallocDefault) != hipSuccess)
            return ICM_ERR_OUT_OF_MEMORY;
        if (hipStreamCreateWithFlags(&own, hipStreamNonBlocking) != hipSuccess) return ICM_ERR_HIP;
        for (auto& e : ev)
            if (hipEventCreate(&e) != hipSuccess) return ICM_ERR_HIP;
        for (auto& e : poll_ev)
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return ICM_ERR_HIP;
        if (hipEventCreate(&ev_sig) != hipSuccess) return ICM_ERR_HIP;
        if (hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming) != hipSuccess) return ICM_ERR_HIP;
        stage_ev.assign(M - 1, nullptr);
        for (auto& e : stage_ev)
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return ICM_ERR_HIP;
        if (hipStreamCreateWithFlags(&aux, hipStreamNonBlocking) != hipSuccess) return ICM_ERR_HIP;
        // ICM_OVERLAP=0 runs the two phases one after the other
        if (const char* env = std::getenv("ICM_OVERLAP")) overlap_on = std::atoi(env) != 0;
        // ICM_SIFT_QUEUE=1: every sift pass from the work queue on a resident
        // grid (else only passes whose one-wave-per-segment grid is too large)
        if (const char* env = std::getenv("ICM_SIFT_QUEUE")) sift_queue = std::atoi(env) != 0;
        // ICM_GRID_LOOP=1: the stage start and the adjoint tile step as
        // grid-stride loops (else only where one wavefront per item would
        // exceed a launch's 2^32 work-items)
        if (const char* env = std::getenv("ICM_GRID_LOOP")) grid_loop = std::atoi(env) != 0;
        tune_init();
        return ICM_OK;
    }

    int alloc_lane(Lane& L, const HostSeries& h)
    {
        int rc;
        const int64_t nt = h.ntiles, ns = (int64_t)h.len.size();
        if ((rc = alloc(&L.hA, h.samples))) return rc;
        if ((rc = alloc(&L.hB, h.samples))) return rc;
        for (int b = 0; b < 2; ++b) {
            if ((rc = alloc(&L.kcnt[b], 2 * nt))) return rc;
            if ((rc = alloc(&L.kpre[b], 2 * nt))) return rc;
            if ((rc = alloc(&L.nk[b], 2 * ns))) return rc;
            if ((rc = alloc(&L.kfix[b], 4 * ns))) return rc;
            if ((rc = alloc(&L.knbx[b], 6 * nt))) return rc;
            if ((rc = alloc(&L.knby[b], 6 * nt))) return rc;
            if ((rc = alloc(&L.kmask[b], 8 * nt))) return rc;
        }
        L.kstride = h.kcap;
        if ((rc = alloc(&L.kc, 2 * h.kcap))) return rc;
        if ((rc = alloc(&L.winlist, (int64_t)kWinItem * h.maxwin))) return rc;
        if ((rc = alloc(&L.wincount, kEpochs))) return rc;
        if ((rc = alloc(&L.ovfcount, kEpochs))) return rc;
        if ((rc = alloc(&L.densecnt, kEpochs))) return rc;
        if ((rc = alloc(&L.squeue, kEpochs))) return rc;
        if ((rc = alloc(&L.ovfsegs, nt))) return rc;
        return ICM_OK;
    }

    PassCtl ctl(const Lane& L, uint32_t e)
    {
        PassCtl pc{};
        pc.epoch = e;
        pc.winlist = L.winlist;
        pc.wincount = L.wincount;
        pc.K = solve_K(cfg.dtype);
        pc.H = solve_H(cfg.dtype);
        pc.err = err;
        pc.dense = L.densecnt;
        return pc;
    }
    KnotsDev<real> knots(const Lane& L, int b)
    {
        KnotsDev<real> k{};
        k.cnt = L.kcnt[b];
        k.pre = L.kpre[b];
        k.n = L.nk[b];
        k.fix = L.kfix[b];
        k.nbx = L.knbx[b];
        k.nby = L.knby[b];
        k.mask = L.kmask[b];
        return k;
    }

    int solve_grid() const { return 4096; }

#define ICM_TRY(x)                                     \
    do {                                               \
        const hipError_t e_ = (x);                     \
        if (e_ != hipSuccess) return hip_fail(e_, __LINE__); \
    } while (0)

    int reset_control(Lane& L, hipStream_t st)
    {
        ICM_TRY(launch_fill_u32((uint32_t*)L.wincount, 0u, kEpochs, st));
        ICM_TRY(launch_fill_u32((uint32_t*)L.ovfcount, 0u, kEpochs, st));
        ICM_TRY(launch_fill_u32((uint32_t*)L.densecnt, 0u, kEpochs, st));
        ICM_TRY(launch_fill_u32((uint32_t*)L.squeue, 0u, kEpochs, st));
        L.epoch = 0;
        return ICM_OK;
    }
    int reset_controls(hipStream_t st)
    {
        int rc;
        if ((rc = reset_control(ln[0], st))) return rc;
        return reset_control(ln[1], st);
    }
    static uint32_t next_epoch(Lane& L) { return ++L.epoch; }

    // Sift passes after the first stage of a phase meet a few tens of knots
    // per 256-sample tile: they run with small knot tables (more waves per
    // SIMD), and the rare tile with more knots with full tables in a second
    // launch, unless more than 1/64 of the input's tiles (counted by the knot
    // prefix of epoch e_in) have more knots: then with full tables.  The
    // result is the same either way.
    void sparse_pass(SiftArgs<real>& a, const Lane& L, const DevSeries& d, uint32_t e_in, uint32_t e)
    {
        a.ovf_segs = L.ovfsegs;
        a.ovf_count = L.ovfcount + e;
        a.dense = L.densecnt + e_in;
        a.dense_max = sparse_dense_max >= 0
                          ? (unsigned)std::min<int64_t>(sparse_dense_max, 0xFFFFFFFFll)
                          : (unsigned)(d.ntiles / 64);
    }

    // The sift launch of pass `pass` of a stage: the segment list and knot
    // tables of the stage's tuning candidate (or, while tuning, of the
    // candidate whose turn it is, timed with events), the pass's work queue.
    int sift_launch(SiftArgs<real>& a, int phase, int stage, int pass, const DevSeries& d,
                    uint32_t e_in, uint32_t e, int64_t bytes, hipStream_t st)
    {
        const int cand = tune_candidate(phase, stage, pass);
        const int l = cand < 0 ? 0 : 1 + cand / 2;
        const bool small = cand < 0 ? stage >= sparse_stage : (cand & 1) != 0;
        const Lane& L = ln[phase];
        a.segs = d.segs[l];
        // one wavefront per segment unless that grid would exceed a launch's
        // 2^32 work-items: then the work-queue instance on a resident grid
        const bool queue = sift_queue || (int64_t)d.nsegs[l] * 64 >= ((int64_t)1 << 32);
        a.queue = queue ? L.squeue + e : nullptr;
        if (small) sparse_pass(a, L, d, e_in, e);
        int rc;
        if ((rc = prof_begin(2, bytes, st))) return rc;
        if ((rc = tune_begin(phase, stage, cand, st))) return rc;
        ICM_TRY(launch_sift<real>(a, d.nsegs[l], st));
        if ((rc = tune_end(cand, st))) return rc;
        return prof_end(st);
    }

    int trace_copy(int phase, int stage, int pass, const DevSeries& d, int b, hipStream_t st)
    {
        if (!trace_passes || pass >= trace_passes || !trace[phase]) return ICM_OK;
        const int64_t n = 2 * (int64_t)d.nseries;
        int32_t* dst = trace[phase] + ((int64_t)stage * trace_passes + pass) * n;
        ICM_TRY(launch_copy<float>((float*)dst, (const float*)ln[phase].nk[b], n, st));
        uint64_t* dh = trace_hash[phase] + ((int64_t)stage * trace_passes + pass) * d.nseries;
        ICM_TRY(launch_knot_digest(d.tile_series, d.tile0, ln[phase].kmask[b], d.ntiles, dh,
                                   d.nseries, st));
        return ICM_OK;
    }

    // Sift one stage of a phase, FIXED mode: init + S x (solve, sift).
    // The stage output (h after S subtractions) ends in `out` if given
    // (series offsets), else in buffer S % 2.
    // h0: the buffer that holds the stage's first pass input when the stage
    // start writes no samples to L.hA (INIT_KNOTS, or INIT_RESID with no
    // h_out: the first pass reads the noise residual itself); null = L.hA.
    int stage_fixed(int phase, int stage, const DevSeries& d, const InitArgs<real>& ia, int S,
                    real* out, hipStream_t st, Rec* record = nullptr, bool full = false,
                    real* h0 = nullptr)
    {
        Lane& L = ln[phase];
        const SeriesDev sd = d.view(L.kstride, cap);
        int rc;
        if (L.epoch + (uint32_t)S + 8 >= kEpochs - 1 && (rc = reset_control(L, st))) return rc;
        const int64_t pass_bytes = 2 * (int64_t)sizeof(real) * d.samples;
        real* in0 = h0 ? h0 : L.hA;
        uint32_t e_emit = next_epoch(L);
        if ((rc = prof_begin(0, pass_bytes, st))) return rc;
        ICM_TRY(launch_init<real>(sd, ia, h0 ? nullptr : L.hA, knots(L, 0), ctl(L, e_emit), st,
                                  grid_loop));
        if ((rc = prof_end(st))) return rc;
        ICM_TRY(launch_knot_prefix<real>(sd, knots(L, 0), in0, ctl(L, e_emit), nullptr, st));
        if ((rc = trace_copy(phase, stage, 0, d, 0, st))) return rc;
        for (int j = 0; j < S; ++j) {
            const int b = j & 1;
            SolveArgs<real> sa{};
            sa.sd = sd;
            sa.k = knots(L, b);
            sa.h = j == 0 ? in0 : (b ? L.hB : L.hA);
            sa.c = L.kc;
            sa.winlist = L.winlist;
            sa.wincount = L.wincount + e_emit;
            sa.active = nullptr;
            sa.err = err;
            sa.nak = cfg.boundary == ICM_BOUNDARY_NOT_A_KNOT;
            if (record) {
                // structure of pass j's input for the adjoint
                const Rec& R = record[j];
                sa.fixf = R.fixf;
                ICM_TRY(launch_copy<float>((float*)R.mask, (const float*)L.kmask[b], 16 * (int64_t)d.ntiles, st));
                ICM_TRY(launch_rec_left(L.knbx[b], R.left, R.right, d.ntiles, st));
                ICM_TRY(launch_copy<float>((float*)R.cnt, (const float*)L.kcnt[b], 2 * (int64_t)d.ntiles, st));
                ICM_TRY(launch_copy<float>((float*)R.pre, (const float*)L.kpre[b], 2 * (int64_t)d.ntiles, st));
                ICM_TRY(launch_copy<float>((float*)R.n, (const float*)L.nk[b], 2 * (int64_t)d.nseries, st));
                ICM_TRY(launch_copy<float>((float*)R.winlist, (const float*)L.winlist,
                                           (int64_t)kWinItem * d.maxwin, st));
                ICM_TRY(launch_copy<float>((float*)R.wincount, (const float*)(L.wincount + e_emit), 1, st));
            }
            if ((rc = prof_begin(1, 0, st))) return rc;
            ICM_TRY(launch_solve<real>(sa, solve_grid(), st));
            if ((rc = prof_end(st))) return rc;
            // a re-sift for the adjoint does not need the stage output itself
            if (record && !full && j == S - 1) break;
            SiftArgs<real> a{};
            a.sd = sd;
            a.h_in = j == 0 ? in0 : (b ? L.hB : L.hA);
            a.h_out = (j == S - 1 && out) ? out : (b ? L.hA : L.hB);
            a.kin = knots(L, b);
            a.c = L.kc;
            a.kout = knots(L, b ^ 1);
            a.emit = (j < S - 1) ? 1 : 0;
            const uint32_t e_in = e_emit;
            e_emit = next_epoch(L);
            a.pc = ctl(L, e_emit);
            a.active = nullptr;
            a.sx_cnt = nullptr;
            a.sx_mark = nullptr;
            if ((rc = sift_launch(a, phase, stage, j, d, e_in, e_emit, pass_bytes, st))) return rc;
            if (j < S - 1) {
                ICM_TRY(launch_knot_prefix<real>(sd, knots(L, b ^ 1), a.h_out, ctl(L, e_emit), nullptr, st));
                if ((rc = trace_copy(phase, stage, j + 1, d, b ^ 1, st))) return rc;
            }
        }
        return ICM_OK;
    }

    // RILLING mode: passes until every series stopped or the cap is reached.
    // Returns with stop_pass[s] = pass whose input h is the exported mode
    // (buffer stop_pass & 1).
    // The passes' kernels return at once when no series of the phase is left
    // (n_active[(j + 1) & 1], the count after pass j - 1, is 0), so the pass
    // loop needs no answer from the device: it is captured whole into the
    // phase's graph when the caps are small (graph_rilling), and otherwise
    // launched with one look at the count per kRillPoll passes, a batch
    // behind (the host reads the count of the previous batch while the
    // current one runs; the batch launched after the last series stopped
    // returns at once).
    static constexpr int kRillPoll = 8;
    int stage_rilling(int phase, int stage, const DevSeries& d, const InitArgs<real>& ia, int max_pass,
                      hipStream_t st)
    {
        Lane& L = ln[phase];
        const SeriesDev sd = d.view(L.kstride, cap);
        ICM_TRY(launch_fill_u32((uint32_t*)active, 1u, d.nseries, st));
        ICM_TRY(launch_fill_u32((uint32_t*)subs, 0u, d.nseries, st));
        ICM_TRY(launch_fill_u32((uint32_t*)stop_pass, 0u, d.nseries, st));
        // before pass 0 every series sifts
        ICM_TRY(launch_fill_u32((uint32_t*)n_active + 1, (uint32_t)d.nseries, 1, st));
        if (L.epoch + 2 * (uint32_t)max_pass + 8 >= kEpochs - 1) {
            int rc = reset_control(L, st);
            if (rc) return rc;
        }
        uint32_t e_emit = next_epoch(L);
        ICM_TRY(launch_init<real>(sd, ia, L.hA, knots(L, 0), ctl(L, e_emit), st, grid_loop));
        ICM_TRY(launch_knot_prefix<real>(sd, knots(L, 0), L.hA, ctl(L, e_emit), nullptr, st));
        int rc;
        if ((rc = trace_copy(phase, stage, 0, d, 0, st))) return rc;
        int pending = -1;   // poll slot whose count the host has not read yet
        for (int j = 0; j < max_pass; ++j) {
            const int b = j & 1;
            const unsigned* run_prev = n_active + ((j + 1) & 1);
            unsigned* run_this = n_active + (j & 1);
            SolveArgs<real> sa{};
            sa.sd = sd;
            sa.k = knots(L, b);
            sa.h = b ? L.hB : L.hA;
            sa.c = L.kc;
            sa.winlist = L.winlist;
            sa.wincount = L.wincount + e_emit;
            sa.active = active;
            sa.err = err;
            sa.nak = cfg.boundary == ICM_BOUNDARY_NOT_A_KNOT;
            sa.run = run_prev;
            ICM_TRY(launch_solve<real>(sa, solve_grid(), st));
            SiftArgs<real> a{};
            a.sd = sd;
            a.h_in = b ? L.hB : L.hA;
            a.h_out = b ? L.hA : L.hB;
            a.kin = knots(L, b);
            a.c = L.kc;
            a.kout = knots(L, b ^ 1);
            a.emit = 1;
            const uint32_t e_in = e_emit;
            e_emit = next_epoch(L);
            a.pc = ctl(L, e_emit);
            a.pc.run = run_prev;
            a.active = active;
            a.sx_cnt = sx_cnt;
            a.sx_mark = sx_mark;
            if ((rc = sift_launch(a, phase, stage, j, d, e_in, e_emit,
                                  2 * (int64_t)sizeof(real) * d.samples, st)))
                return rc;
            ICM_TRY(launch_fill_u32((uint32_t*)run_this, 0u, 1, st));
            ICM_TRY(launch_rilling_decide<real>(sd, L.nk[b], sx_cnt, sx_mark, active, stop_pass, subs,
                                                j, max_pass, run_prev, run_this, st));
            PassCtl pn = ctl(L, e_emit);
            pn.run = run_this;
            ICM_TRY(launch_knot_prefix<real>(sd, knots(L, b ^ 1), a.h_out, pn, active, st));
            if ((rc = trace_copy(phase, stage, j + 1, d, b ^ 1, st))) return rc;
            // early exit: results are identical to running every pass to the
            // cap with finished series masked (they never change again).
            // Not while a graph is being captured: it holds every pass.
            if (!capturing && ((j + 1) % kRillPoll == 0 || j == max_pass - 1)) {
                const int slot = (j / kRillPoll) & 1;
                ICM_TRY(hipMemcpyAsync(h_n_active + slot, run_this, sizeof(unsigned),
                                       hipMemcpyDeviceToHost, st));
                ICM_TRY(hipEventRecord(poll_ev[slot], st));
                if (pending >= 0) {
                    ICM_TRY(hipEventSynchronize(poll_ev[pending]));
                    if (h_n_active[pending] == 0) break;
                }
                pending = slot;
            }
        }
        if (stops[phase]) {
            int32_t* dst = stops[phase] + (int64_t)stage * d.nseries;
            ICM_TRY(launch_copy<float>((float*)dst, (const float*)subs, d.nseries, st));
        }
        return ICM_OK;
    }

    // Noise phase: E_{k+1}(w^i) for k = 0..M-2 (cudaICEEMDAN.cu:845-994).
    // With `done`, the event done[k] is recorded after stage k.
    int noise_phase(hipStream_t st, hipEvent_t* done = nullptr)
    {
        const int M = cfg.num_imfs;
        int rc;
        // stage k starts from the residual w - E_0 - ... - E_{k-1}; stage
        // k >= 1 forms it from the previous residual and E_{k-1} while it
        // reads it (k_init_w, INIT_RESID), alternating two buffers
        // FIXED: the stage's first pass reads the residual buffer itself, so
        // the stage start writes no copy of it (RILLING keeps the copy in
        // the sample buffer: a series that stops at once exports it)
        real* res[2] = {wres, wres2};
        const bool fixed = cfg.stop_mode == ICM_STOP_FIXED;
        for (int k = 0; k < M - 1; ++k) {
            InitArgs<real> ia{};
            if (k == 0) {
                ia.mode = fixed ? INIT_KNOTS : INIT_COPY;
                ia.src = wres;
            } else {
                ia.mode = INIT_RESID;
                ia.src = res[(k - 1) & 1];
                ia.E = modes + (int64_t)(k - 1) * hn.samples;
                ia.resid = res[k & 1];
            }
            real* Ek = modes + (int64_t)k * hn.samples;
            if (fixed) {
                if ((rc = stage_fixed(ICM_PHASE_NOISE, k, dn, ia, cfg.noise_max_sifts, Ek, st,
                                      nullptr, false, k == 0 ? wres : res[k & 1])))
                    return rc;
            } else {
                if ((rc = stage_rilling(ICM_PHASE_NOISE, k, dn, ia, cfg.noise_max_sifts, st)))
                    return rc;
                ICM_TRY(launch_select_copy<real>(dn.view(ln[0].kstride, cap), dn.max_len, ln[0].hA,
                                                 ln[0].hB, stop_pass,
                                                 Ek, st));
            }
            if (done) ICM_TRY(hipEventRecord(done[k], st));
        }
        return ICM_OK;
    }

    // CEEMDAN stages (cudaICEEMDAN.cu:1003-1298)
    // stage-0 statistics: var(x) per channel, var(E_1(w^i)) per noise series
    int signal_begin(hipStream_t st)
    {
        const int C = cfg.num_channels;
        ICM_TRY(launch_series_var<real>(y, d_chan_off, 0, d_chan_len, C, var_chan, st));
        for (int c = 0; c < C; ++c)
            ICM_TRY(launch_standardize<real>(y + chan_off[c], var_chan + c, xt + chan_off[c],
                                             lens[c], st));
        ICM_TRY(launch_series_var<real>(modes, dn.off, 0, dn.len, dn.nseries, var_noise, st));
        ICM_TRY(launch_inv_sqrt<real>(var_noise, inv_noise, dn.nseries, 0, st));
        return ICM_OK;
    }

    // CEEMDAN stage k: r_k -> r_{k+1}, IMF_k (cudaICEEMDAN.cu:1003-1298)
    int signal_stage(int k, hipStream_t st)
    {
        const int M = cfg.num_imfs, I = cfg.num_noise, C = cfg.num_channels;
        const real eps0 = (real)cfg.noise_strength;
        int rc;
        const real* Ek = modes + (int64_t)k * hn.samples;
        InitArgs<real> ia{};
        ia.E = Ek;
        ia.eps0 = eps0;
        ia.chan_off = d_chan_off;
        if (k == 0) {
            ia.mode = INIT_STAGE0;
            ia.chan_sig = xt;
            ia.scale = inv_noise;
        } else {
            ICM_TRY(launch_series_var<real>(r, d_chan_off, 0, d_chan_len, C, var_chan, st));
            ICM_TRY(launch_inv_sqrt<real>(var_chan, sq_chan, C, 1, st));
            ia.mode = INIT_STAGEK;
            ia.chan_sig = r;
            ia.scale = sq_chan;
        }
        const int* parity = nullptr;
        int fixed_parity = 0;
        if (cfg.stop_mode == ICM_STOP_FIXED) {
            Rec* R = krec.empty() ? nullptr : krec[k].data();
            if ((rc = stage_fixed(ICM_PHASE_SIGNAL, k, dsig, ia, cfg.max_sifts, nullptr, st, R,
                                  true)))
                return rc;
            if (R) krec_valid[k] = 1;
            fixed_parity = cfg.max_sifts & 1;
        } else {
            if ((rc = stage_rilling(ICM_PHASE_SIGNAL, k, dsig, ia, cfg.max_sifts, st))) return rc;
            parity = stop_pass;
        }
        ICM_TRY(launch_ensemble<real>(k, C, I, max_chan_len, d_chan_len, d_chan_off, dsig.off,
                                      dsig.eoff, dsig.nidx, ln[1].hA, ln[1].hB, parity,
                                      fixed_parity, Ek,
                                      inv_noise, sq_chan, eps0, xt, r, imf, d_chan_imf_off, st));
        // residues kept for the adjoint (r_0 = x~ is xt)
        if (k + 1 <= M - 2)
            ICM_TRY(launch_copy<real>(r_hist + (int64_t)(k + 1) * chan_total, r, chan_total, st));
        return ICM_OK;
    }

    // residue row M-1 (cudaICEEMDAN.cu:1298)
    int signal_end(hipStream_t st)
    {
        const int M = cfg.num_imfs, C = cfg.num_channels;
        for (int c = 0; c < C; ++c)
            ICM_TRY(launch_copy<real>(imf + chan_imf_off[c] + (int64_t)(M - 1) * lens[c],
                                      r + chan_off[c], lens[c], st));
        return ICM_OK;
    }

    // With `ready`, the stage-0 statistics wait for ready[0] and stage k for
    // ready[k] (the noise mode E_{k+1} it reads).
    int signal_phase(hipStream_t st, hipEvent_t* ready = nullptr)
    {
        int rc;
        if (ready) ICM_TRY(hipStreamWaitEvent(st, ready[0], 0));
        if ((rc = signal_begin(st))) return rc;
        for (int k = 0; k < cfg.num_imfs - 1; ++k) {
            if (ready && k > 0) ICM_TRY(hipStreamWaitEvent(st, ready[k], 0));
            if ((rc = signal_stage(k, st))) return rc;
        }
        return signal_end(st);
    }

    // Both phases side by side (FIXED mode): the noise-mode phase on the
    // plan's second stream, the CEEMDAN phase on `st`; CEEMDAN stage k waits
    // only for noise-mode stage k (it reads E_{k+1}), so the latency-bound
    // launches of one phase (knot prefix, solve, stage start, ensemble) run
    // beside the other's sift passes.  The result is the same bits as one
    // phase after the other: the phases share no buffer but the noise modes,
    // which the CEEMDAN phase reads after the event that follows their write.
    bool overlap() const
    {
        return cfg.stop_mode == ICM_STOP_FIXED && overlap_on && aux && !tuning && !trace_passes;
    }
    int both_phases(hipStream_t st)
    {
        int rc;
        ICM_TRY(hipEventRecord(ev_fork, st));
        ICM_TRY(hipStreamWaitEvent(aux, ev_fork, 0));
        if ((rc = noise_phase(aux, stage_ev.data()))) return rc;
        if (!capturing) ICM_TRY(hipEventRecord(ev[1], aux));
        if (!capturing) ICM_TRY(hipEventRecord(ev_sig, st));
        return signal_phase(st, stage_ev.data());
    }

    // FIXED: always; RILLING: when both caps are at most kRillGraphCap passes
    // (the graph then holds every pass up to the caps; passes after the last
    // series stopped return at once)
    static constexpr int kRillGraphCap = 64;
    bool graph_mode() const
    {
        return cfg.stop_mode == ICM_STOP_FIXED ||
               (cfg.max_sifts <= kRillGraphCap && cfg.noise_max_sifts <= kRillGraphCap);
    }

    bool graph_on() const
    {
        return cfg.use_graph && graph_mode() && !trace_passes && !prof && !graph_failed && !tuning;
    }

    int capture_or_run(int which, hipStream_t st)
    {
        auto body = [&](hipStream_t s) -> int {
            return which == 0 ? noise_phase(s) : which == 1 ? signal_phase(s) : both_phases(s);
        };
        if (!graph_on()) return body(st);
        if (!gexec[which]) {
            // capture on the plan's own stream (and, for both phases, on
            // the second stream that joins it), replay on the caller's
            const uint32_t e0[2] = {ln[0].epoch, ln[1].epoch};
            const hipError_t be = hipStreamBeginCapture(own, hipStreamCaptureModeThreadLocal);
            if (be != hipSuccess) {
                graph_failed = true;
                (void)hip_fail(be, __LINE__);
                return body(st);
            }
            capturing = true;
            int rc = body(own);
            capturing = false;
            hipGraph_t g = nullptr;
            hipError_t ce = hipStreamEndCapture(own, &g);
            if (rc != ICM_OK || ce != hipSuccess || !g) {
                if (g) (void)hipGraphDestroy(g);
                (void)hip_fail(ce != hipSuccess ? ce : hipErrorUnknown, __LINE__);
                (void)hipGetLastError();
                graph_failed = true;
                ln[0].epoch = e0[0];
                ln[1].epoch = e0[1];
                return body(st);
            }
            hipGraphExec_t ex = nullptr;
            const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            if (ie != hipSuccess) {
                (void)hip_fail(ie, __LINE__);
                (void)hipGraphDestroy(g);
                (void)hipGetLastError();
                graph_failed = true;
                ln[0].epoch = e0[0];
                ln[1].epoch = e0[1];
                return body(st);
            }
            (void)hipGraphDestroy(g);
            gexec[which] = ex;
        }
        // the graph's launches carry the epochs assigned at capture; the
        // control words were reset before the replay
        ICM_TRY(hipGraphLaunch(gexec[which], st));
        return ICM_OK;
    }

    int forward(const void* const* sig, void* const* out, const void* noise, hipStream_t st) override
    {
        const int C = cfg.num_channels, M = cfg.num_imfs;
        *h_err = 0;
        stage_ = -1;
        // the saved state is overwritten from here on: a backward is valid
        // again only after this forward has completed
        have_forward = false;
        std::fill(krec_valid.begin(), krec_valid.end(), (char)0);
        for (int c = 0; c < C; ++c)
            if (!out[c]) return ICM_ERR_INVALID_ARG;
        // a tuning forward times the sift candidates (eagerly, with events)
        tuning = tune_left > 0;
        tused = 0;
        ICM_TRY(hipEventRecord(ev[0], st));
        ICM_TRY(launch_fill_u32((uint32_t*)err, 0u, 1, st));
        int rc;
        if ((rc = inputs_and_noise(sig, noise, st))) return rc;
        if ((rc = reset_controls(st))) return rc;
        const bool both = overlap();
        const bool split = !(both && graph_on());   // phase times measured
        if (both) {
            if ((rc = capture_or_run(2, st))) return rc;
        } else {
            if ((rc = capture_or_run(0, st))) return rc;
            ICM_TRY(hipEventRecord(ev[1], st));
            ICM_TRY(hipEventRecord(ev_sig, st));
            if ((rc = capture_or_run(1, st))) return rc;
        }
        ICM_TRY(hipEventRecord(ev[2], st));
        for (int c = 0; c < C; ++c)
            ICM_TRY(launch_copy<real>((real*)out[c], imf + chan_imf_off[c], (int64_t)M * lens[c], st));
        ICM_TRY(hipEventRecord(ev[3], st));
        ICM_TRY(hipMemcpyAsync(h_err, err, sizeof(int), hipMemcpyDeviceToHost, st));
        ICM_TRY(hipStreamSynchronize(st));
        float t01 = 0, t12 = 0, t03 = 0;
        (void)hipEventElapsedTime(&t03, ev[0], ev[3]);
        if (split) {
            (void)hipEventElapsedTime(&t01, ev[0], ev[1]);
            (void)hipEventElapsedTime(&t12, ev_sig, ev[2]);
        }
        timing.noise_ms = split ? t01 : -1.0;
        timing.signal_ms = split ? t12 : -1.0;
        timing.total_ms = t03;
        prof_collect();
        if (*h_err) {
            tuning = false;
            tused = 0;
            return ICM_ERR_DEVICE;
        }
        if (tuning) tune_collect();
        have_forward = true;
        return ICM_OK;
    }

    // inputs into the plan (d_signals may be null: y is then left as it is)
    // and the white noise w^i (cudaICEEMDAN.cu:739-741), one set per
    // distinct length, or the caller's noise
    int inputs_and_noise(const void* const* sig, const void* noise, hipStream_t st)
    {
        const int C = cfg.num_channels;
        if (sig) {
            for (int c = 0; c < C; ++c) {
                if (!sig[c]) return ICM_ERR_INVALID_ARG;
                ICM_TRY(launch_copy<real>(y + chan_off[c], (const real*)sig[c], lens[c], st));
            }
        }
        noise_external = (noise != nullptr);
        if (noise_external) {
            if (glen.size() != 1) return ICM_ERR_INVALID_ARG;
            ICM_TRY(launch_copy<real>(wres, (const real*)noise, hn.samples, st));
        } else {
            for (size_t g = 0; g < glen.size(); ++g)
                ICM_TRY(launch_noise<real>(cfg.seed, cfg.num_noise, glen[g], wres + group_off[g],
                                           glen[g], st));
        }
        return ICM_OK;
    }

    int sync_err(hipStream_t st)
    {
        ICM_TRY(hipMemcpyAsync(h_err, err, sizeof(int), hipMemcpyDeviceToHost, st));
        ICM_TRY(hipStreamSynchronize(st));
        return *h_err ? ICM_ERR_DEVICE : ICM_OK;
    }

    // ---- staged forward: noise modes and stage-0 statistics, then CEEMDAN
    // stages a few at a time, with the state between stages saveable
    int forward_begin(const void* const* sig, const void* noise, hipStream_t st) override
    {
        tuning = false;
        *h_err = 0;
        have_forward = false;
        std::fill(krec_valid.begin(), krec_valid.end(), (char)0);
        ICM_TRY(launch_fill_u32((uint32_t*)err, 0u, 1, st));
        int rc;
        if ((rc = inputs_and_noise(sig, noise, st))) return rc;
        if ((rc = reset_controls(st))) return rc;
        if ((rc = noise_phase(st))) return rc;
        if ((rc = signal_begin(st))) return rc;
        if ((rc = sync_err(st))) return rc;
        stage_ = 0;
        return ICM_OK;
    }

    int forward_stages(int count, int* stage_out, hipStream_t st) override
    {
        if (stage_ < 0) return ICM_ERR_STATE;
        int rc;
        for (int n = 0; n < count && stage_ < cfg.num_imfs - 1; ++n) {
            if ((rc = signal_stage(stage_, st))) return rc;
            ++stage_;
        }
        if ((rc = sync_err(st))) return rc;
        if (stage_out) *stage_out = stage_;
        return ICM_OK;
    }

    int forward_finish(void* const* out, hipStream_t st) override
    {
        const int C = cfg.num_channels, M = cfg.num_imfs;
        if (stage_ != M - 1) return ICM_ERR_STATE;
        for (int c = 0; c < C; ++c)
            if (!out[c]) return ICM_ERR_INVALID_ARG;
        int rc;
        if ((rc = signal_end(st))) return rc;
        for (int c = 0; c < C; ++c)
            ICM_TRY(launch_copy<real>((real*)out[c], imf + chan_imf_off[c], (int64_t)M * lens[c], st));
        if ((rc = sync_err(st))) return rc;
        stage_ = -1;
        have_forward = true;
        return ICM_OK;
    }

    // state between stages: header, then y, r_k, IMF rows, residue history
    struct StateHeader {
        uint64_t magic, version;
        int32_t dtype, num_channels, num_noise, num_imfs, stop_mode, max_sifts, noise_max_sifts,
            noise_external, stage, pad;
        uint64_t seed;
        double noise_strength;
        int64_t chan_total, len_hash;
    };
    static constexpr uint64_t kStateMagic = 0x5445415453434d49ull;   // "IMCSTATE"

    int64_t len_hash() const
    {
        uint64_t h = 1469598103934665603ull;
        for (int64_t L : lens) h = (h ^ (uint64_t)L) * 1099511628211ull;
        return (int64_t)(h >> 1);
    }
    int64_t state_bytes() const override
    {
        const int M = cfg.num_imfs;
        return (int64_t)sizeof(StateHeader) +
               (int64_t)sizeof(real) * chan_total * (1 + 1 + M + (M - 1));
    }
    StateHeader header() const
    {
        StateHeader h{};
        h.magic = kStateMagic;
        h.version = 1;
        h.dtype = cfg.dtype;
        h.num_channels = cfg.num_channels;
        h.num_noise = cfg.num_noise;
        h.num_imfs = cfg.num_imfs;
        h.stop_mode = cfg.stop_mode;
        h.max_sifts = cfg.max_sifts;
        h.noise_max_sifts = cfg.noise_max_sifts;
        h.noise_external = noise_external ? 1 : 0;
        h.stage = stage_;
        h.pad = cfg.boundary;
        h.seed = cfg.seed;
        h.noise_strength = cfg.noise_strength;
        h.chan_total = chan_total;
        h.len_hash = len_hash();
        return h;
    }
    int state_save(void* host, int64_t bytes, hipStream_t st) override
    {
        if (stage_ < 0) return ICM_ERR_STATE;
        if (!host || bytes < state_bytes()) return ICM_ERR_INVALID_ARG;
        const int M = cfg.num_imfs;
        ICM_TRY(hipStreamSynchronize(st));
        const StateHeader h = header();
        std::memcpy(host, &h, sizeof h);
        char* p = (char*)host + sizeof h;
        const size_t row = sizeof(real) * (size_t)chan_total;
        ICM_TRY(hipMemcpy(p, y, row, hipMemcpyDeviceToHost));
        ICM_TRY(hipMemcpy(p + row, r, row, hipMemcpyDeviceToHost));
        ICM_TRY(hipMemcpy(p + 2 * row, imf, row * M, hipMemcpyDeviceToHost));
        ICM_TRY(hipMemcpy(p + (2 + M) * row, r_hist, row * (M - 1), hipMemcpyDeviceToHost));
        return ICM_OK;
    }
    int state_load(const void* host, int64_t bytes, const void* noise, hipStream_t st) override
    {
        tuning = false;
        if (!host || bytes < state_bytes()) return ICM_ERR_INVALID_ARG;
        StateHeader h;
        std::memcpy(&h, host, sizeof h);
        StateHeader mine = header();
        if (h.magic != kStateMagic || h.version != 1 || h.dtype != mine.dtype ||
            h.num_channels != mine.num_channels || h.num_noise != mine.num_noise ||
            h.num_imfs != mine.num_imfs || h.stop_mode != mine.stop_mode ||
            h.max_sifts != mine.max_sifts || h.noise_max_sifts != mine.noise_max_sifts ||
            h.seed != mine.seed || h.noise_strength != mine.noise_strength || h.pad != mine.pad ||
            h.chan_total != mine.chan_total || h.len_hash != mine.len_hash ||
            h.stage < 0 || h.stage > cfg.num_imfs - 1)
            return ICM_ERR_INVALID_ARG;
        // the noise modes are recomputed (deterministic), from the same noise
        if ((h.noise_external != 0) != (noise != nullptr)) return ICM_ERR_INVALID_ARG;
        const int M = cfg.num_imfs;
        const char* p = (const char*)host + sizeof h;
        const size_t row = sizeof(real) * (size_t)chan_total;
        ICM_TRY(hipStreamSynchronize(st));
        ICM_TRY(hipMemcpy(y, p, row, hipMemcpyHostToDevice));
        int rc;
        *h_err = 0;
        have_forward = false;
        std::fill(krec_valid.begin(), krec_valid.end(), (char)0);
        ICM_TRY(launch_fill_u32((uint32_t*)err, 0u, 1, st));
        if ((rc = inputs_and_noise(nullptr, noise, st))) return rc;
        if ((rc = reset_controls(st))) return rc;
        if ((rc = noise_phase(st))) return rc;
        if ((rc = signal_begin(st))) return rc;
        if ((rc = sync_err(st))) return rc;
        ICM_TRY(hipMemcpy(r, p + row, row, hipMemcpyHostToDevice));
        ICM_TRY(hipMemcpy(imf, p + 2 * row, row * M, hipMemcpyHostToDevice));
        ICM_TRY(hipMemcpy(r_hist, p + (2 + M) * row, row * (M - 1), hipMemcpyHostToDevice));
        stage_ = h.stage;
        return ICM_OK;
    }

    int alloc_recs(std::vector<Rec>& out)
    {
        const int S = cfg.max_sifts;
        int rc;
        std::vector<Rec> rr(S);
        for (int j = 0; j < S; ++j) {
            if ((rc = alloc(&rr[j].mask, 8 * (int64_t)dsig.ntiles))) return rc;
            if ((rc = alloc(&rr[j].left, 2 * (int64_t)dsig.ntiles))) return rc;
            if ((rc = alloc(&rr[j].right, 2 * (int64_t)dsig.ntiles))) return rc;
            if ((rc = alloc(&rr[j].cnt, 2 * (int64_t)dsig.ntiles))) return rc;
            if ((rc = alloc(&rr[j].pre, 2 * (int64_t)dsig.ntiles))) return rc;
            if ((rc = alloc(&rr[j].n, 2 * (int64_t)dsig.nseries))) return rc;
            if ((rc = alloc(&rr[j].fixf, 4 * (int64_t)dsig.nseries))) return rc;
            if ((rc = alloc(&rr[j].winlist, (int64_t)kWinItem * dsig.maxwin))) return rc;
            if ((rc = alloc(&rr[j].wincount, 1))) return rc;
        }
        out.swap(rr);
        return ICM_OK;
    }

    int backward_alloc()
    {
        if (mu) return ICM_OK;
        const int M = cfg.num_imfs;
        int rc;
        if ((rc = alloc(&mu, dsig.samples))) return rc;
        if ((rc = alloc(&mom, 8 * ln[ICM_PHASE_SIGNAL].kstride))) return rc;
        if ((rc = alloc(&edge, 16 * (int64_t)dsig.ntiles))) return rc;
        if ((rc = alloc(&ga, 2 * ln[ICM_PHASE_SIGNAL].kstride))) return rc;
        if ((rc = alloc(&gr, (int64_t)M * chan_total))) return rc;
        if ((rc = alloc(&gacc, chan_total))) return rc;
        if ((rc = alloc(&dots, (int64_t)dsig.nseries))) return rc;
        if ((rc = alloc(&gy, chan_total))) return rc;
        return ICM_OK;
    }

    // Reverse mode over the signal phase of the last forward.
    int backward(const void* const* gimf, void* const* gsig, hipStream_t st)
    {
        tuning = false;
        if (!have_forward || cfg.stop_mode != ICM_STOP_FIXED ||
            cfg.boundary != ICM_BOUNDARY_NATURAL)
            return ICM_ERR_STATE;
        const int C = cfg.num_channels, M = cfg.num_imfs, I = cfg.num_noise, S = cfg.max_sifts;
        const real eps0 = (real)cfg.noise_strength;
        int rc;
        if ((rc = backward_alloc())) return rc;
        *h_err = 0;
        ICM_TRY(hipEventRecord(ev[0], st));
        ICM_TRY(launch_fill_u32((uint32_t*)err, 0u, 1, st));
        // dL/dr_k: IMF_0 = x~ - r_1, IMF_k = r_k - r_{k+1}, IMF_{M-1} = r_{M-1}
        for (int c = 0; c < C; ++c) {
            if (!gimf[c] || !gsig[c]) return ICM_ERR_INVALID_ARG;
            ICM_TRY(launch_adj_seed<real>((const real*)gimf[c], M, lens[c], chan_total,
                                          gr + chan_off[c], st));
        }
        if ((rc = reset_control(ln[ICM_PHASE_SIGNAL], st))) return rc;
        const SeriesDev sd = dsig.view(ln[ICM_PHASE_SIGNAL].kstride, cap);
        for (int k = M - 2; k >= 0; --k) {
            const real* Ek = modes + (int64_t)k * hn.samples;
            const real* rk = k == 0 ? xt : r_hist + (int64_t)k * chan_total;
            InitArgs<real> ia{};
            ia.E = Ek;
            ia.eps0 = eps0;
            ia.chan_off = d_chan_off;
            if (k == 0) {
                ia.mode = INIT_STAGE0;
                ia.chan_sig = xt;
                ia.scale = inv_noise;
            } else {
                ICM_TRY(launch_series_var<real>(rk, d_chan_off, 0, d_chan_len, C, var_chan, st));
                ICM_TRY(launch_inv_sqrt<real>(var_chan, sq_chan, C, 1, st));
                ia.mode = INIT_STAGEK;
                ia.chan_sig = rk;
                ia.scale = sq_chan;
            }
            Rec* R = nullptr;
            if (!krec.empty() && krec_valid[k]) {
                R = krec[k].data();   // kept by the forward
            } else {
                if (rec.empty() && (rc = alloc_recs(rec))) return rc;
                R = rec.data();
                if ((rc = stage_fixed(ICM_PHASE_SIGNAL, k, dsig, ia, S, nullptr, st, R))) return rc;
            }
            const real* lam = gr + (int64_t)(k + 1) * chan_total;
            auto rec_of = [&](int j) {
                PassRec r{};
                if (j >= 0 && j < S) {
                    r.mask = R[j].mask;
                    r.left = R[j].left;
                    r.right = R[j].right;
                    r.cnt = R[j].cnt;
                    r.pre = R[j].pre;
                    r.n = R[j].n;
                    r.fixf = R[j].fixf;
                }
                return r;
            };
            auto adj_args = [&](int j) {
                AdjArgs<real> a{};
                a.sd = sd;
                a.ps = rec_of(j);
                a.mu = mu;
                a.mom = mom;
                a.edge = edge;
                a.ga = ga;
                a.winlist = R[j].winlist;
                a.wincount = R[j].wincount;
                a.err = err;
                return a;
            };
            // fused per-tile steps: (broadcast | scatter of pass j+1) + moments of pass j
            for (int j = S - 1; j >= -1; --j) {
                AdjTileArgs<real> ta{};
                ta.sd = sd;
                ta.mu = mu;
                ta.bcast = j == S - 1 ? lam : nullptr;
                ta.chan_off = d_chan_off;
                ta.sc = rec_of(j + 1);
                ta.ga = ga;
                ta.mo = rec_of(j);
                ta.mom = mom;
                ta.edge = edge;
                ICM_TRY(launch_adj_tile<real>(ta, st, grid_loop));
                if (j >= 0) ICM_TRY(launch_adj_span_knot<real>(adj_args(j), solve_grid(), st));
            }
            ICM_TRY(launch_adj_stage<real>(k, C, I, max_chan_len, d_chan_len, d_chan_off, sd, lam,
                                           mu, Ek, rk, eps0, gr + (int64_t)k * chan_total, gacc,
                                           dots, st));
        }
        ICM_TRY(launch_adj_input<real>(C, d_chan_len, d_chan_off, y, gr, gy, st));
        for (int c = 0; c < C; ++c)
            ICM_TRY(launch_copy<real>((real*)gsig[c], gy + chan_off[c], lens[c], st));
        ICM_TRY(hipEventRecord(ev[3], st));
        ICM_TRY(hipMemcpyAsync(h_err, err, sizeof(int), hipMemcpyDeviceToHost, st));
        ICM_TRY(hipStreamSynchronize(st));
        float t03 = 0;
        (void)hipEventElapsedTime(&t03, ev[0], ev[3]);
        timing.backward_ms = t03;
        if (*h_err) return ICM_ERR_DEVICE;
        return ICM_OK;
    }

    // one pass over the noise-phase series (I series of length N, back to back)
    int sift_once(const real* h, real* hn, int32_t* nk_out, hipStream_t st)
    {
        tuning = false;
        int rc;
        ICM_TRY(launch_fill_u32((uint32_t*)err, 0u, 1, st));
        if ((rc = reset_control(ln[ICM_PHASE_NOISE], st))) return rc;
        InitArgs<real> ia{};
        ia.mode = INIT_COPY;
        ia.src = h;
        if ((rc = stage_fixed(ICM_PHASE_NOISE, 0, dn, ia, 1, hn, st))) return rc;
        if (nk_out)
            ICM_TRY(launch_copy<float>((float*)nk_out, (const float*)ln[ICM_PHASE_NOISE].nk[0],
                                       2 * dn.nseries, st));
        ICM_TRY(hipMemcpyAsync(h_err, err, sizeof(int), hipMemcpyDeviceToHost, st));
        ICM_TRY(hipStreamSynchronize(st));
        return *h_err ? ICM_ERR_DEVICE : ICM_OK;
    }

    int noise_modes_out(void* d_out, int64_t count, hipStream_t st) override
    {
        const int64_t n = (int64_t)(cfg.num_imfs - 1) * hn.samples;
        if (count < n) return ICM_ERR_INVALID_ARG;
        ICM_TRY(launch_copy<real>((real*)d_out, modes, n, st));
        ICM_TRY(hipStreamSynchronize(st));
        return ICM_OK;
    }
#undef ICM_TRY
};

}  // namespace icm

using namespace icm;

struct icm_plan_s {
    PlanBase* p;
};

extern "C" {

const char* icm_last_error(void) { return g_last_error; }

const char* icm_status_string(int s)
{
    switch (s) {
    case ICM_OK: return "ok";
    case ICM_ERR_INVALID_ARG: return "invalid argument";
    case ICM_ERR_OUT_OF_MEMORY: return "out of device memory";
    case ICM_ERR_HIP: return "HIP runtime error";
    case ICM_ERR_DEVICE: return "device kernel error";
    case ICM_ERR_INDEX: return "sample-index array is not 0..N-1";
    case ICM_ERR_STATE: return "call not valid in this plan state";
    default: return "unknown status";
    }
}

int icm_config_default(icm_config* c)
{
    if (!c) return ICM_ERR_INVALID_ARG;
    std::memset(c, 0, sizeof(*c));
    c->dtype = ICM_DTYPE_F64;
    c->num_channels = 1;
    c->num_noise = 100;
    c->num_imfs = 10;
    c->stop_mode = ICM_STOP_FIXED;
    c->max_sifts = 10;
    c->noise_max_sifts = 10;
    c->noise_strength = 0.2;
    c->seed = 1234;
    c->use_graph = 1;
    return ICM_OK;
}

int icm_validate_config(const icm_config* c) { return validate(c); }

int icm_plan_bytes(const icm_config* c, int64_t* bytes)
{
    int rc = validate(c);
    if (rc) return rc;
    if (!bytes) return ICM_ERR_INVALID_ARG;
    *bytes = plan_bytes_estimate(c);
    return ICM_OK;
}

int icm_plan_create(const icm_config* c, icm_plan* out)
{
    if (!out) return ICM_ERR_INVALID_ARG;
    *out = nullptr;
    int rc = validate(c);
    if (rc) return rc;
    // refuse a plan that cannot fit before allocating any of it: against
    // ICM_MEMORY_BUDGET (bytes) when set, else the device's free memory
    const int64_t need = plan_bytes_estimate(c);
    int64_t budget = -1;
    const char* what = "budget";
    if (const char* env = std::getenv("ICM_MEMORY_BUDGET")) {
        budget = std::atoll(env);
    } else {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) == hipSuccess) {
            budget = (int64_t)fr;
            what = "free device memory";
        }
    }
    if (budget >= 0 && need > budget) {
        std::snprintf(g_last_error, sizeof g_last_error,
                      "plan needs %lld bytes, %s is %lld bytes", (long long)need, what,
                      (long long)budget);
        return ICM_ERR_OUT_OF_MEMORY;
    }
    PlanBase* p = nullptr;
    if (c->dtype == ICM_DTYPE_F64) p = new (std::nothrow) Plan<double>();
    else p = new (std::nothrow) Plan<float>();
    if (!p) return ICM_ERR_OUT_OF_MEMORY;
    p->cfg = *c;
    layout(*p);               // copies the lengths; the caller's array is not kept
    p->cfg.lengths = nullptr;
    if (c->dtype == ICM_DTYPE_F64) rc = static_cast<Plan<double>*>(p)->setup();
    else rc = static_cast<Plan<float>*>(p)->setup();
    if (rc) {
        p->free_all();
        delete p;
        return rc;
    }
    icm_plan h = new (std::nothrow) icm_plan_s;
    if (!h) {
        p->free_all();
        delete p;
        return ICM_ERR_OUT_OF_MEMORY;
    }
    h->p = p;
    *out = h;
    return ICM_OK;
}

int icm_plan_destroy(icm_plan plan)
{
    if (!plan) return ICM_ERR_INVALID_ARG;
    // a plan destroyed before its last tuning forward leaves what it measured
    if (plan->p->tune_left > 0 && plan->p->tune_round > 0) plan->p->tune_finish();
    plan->p->free_all();
    delete plan->p;
    delete plan;
    return ICM_OK;
}

int icm_plan_get_info(icm_plan plan, icm_plan_info* info)
{
    if (!plan || !info) return ICM_ERR_INVALID_ARG;
    PlanBase& P = *plan->p;
    info->device_bytes = P.bytes;
    info->noise_series = (int64_t)P.hn.len.size();
    info->signal_series = (int64_t)P.hsig.len.size();
    info->noise_samples = P.hn.samples;
    info->signal_samples = P.hsig.samples;
    info->tiles_noise = P.hn.ntiles;
    info->tiles_signal = P.hsig.ntiles;
    info->tile = 2 * (P.cap - 2);
    info->solve_rows_per_lane = solve_B(P.cfg.dtype);
    info->solve_halo = solve_H(P.cfg.dtype);
    return ICM_OK;
}

int icm_forward(icm_plan plan, const void* const* d_signals, void* const* d_imfs,
                const void* d_noise, void* stream)
{
    if (!plan || !d_signals || !d_imfs) return ICM_ERR_INVALID_ARG;
    return plan->p->forward(d_signals, d_imfs, d_noise, (hipStream_t)stream);
}

int icm_last_timing(icm_plan plan, icm_timing* out)
{
    if (!plan || !out) return ICM_ERR_INVALID_ARG;
    *out = plan->p->timing;
    return ICM_OK;
}

int icm_set_profiling(icm_plan plan, int on)
{
    if (!plan) return ICM_ERR_INVALID_ARG;
    plan->p->prof = on ? 1 : 0;
    return ICM_OK;
}

int icm_get_kernel_stats(icm_plan plan, int kind, double* total_ms, int64_t* launches,
                         int64_t* alg_bytes)
{
    if (!plan || kind < 0 || kind > 2 || !total_ms || !launches || !alg_bytes)
        return ICM_ERR_INVALID_ARG;
    *total_ms = plan->p->kern_ms[kind];
    *launches = plan->p->kern_n[kind];
    *alg_bytes = plan->p->kern_bytes[kind];
    return ICM_OK;
}

int icm_set_autotune(icm_plan plan, int forwards)
{
    if (!plan) return ICM_ERR_INVALID_ARG;
    return plan->p->set_autotune(forwards);
}

int icm_get_tuning(icm_plan plan, int phase, int32_t* host_out, int64_t count)
{
    if (!plan) return ICM_ERR_INVALID_ARG;
    return plan->p->get_tuning(phase, host_out, count);
}

int icm_set_tuning(icm_plan plan, int phase, const int32_t* host_in, int64_t count)
{
    if (!plan) return ICM_ERR_INVALID_ARG;
    return plan->p->set_tuning(phase, host_in, count);
}

int icm_set_overlap(icm_plan plan, int on)
{
    if (!plan) return ICM_ERR_INVALID_ARG;
    plan->p->overlap_on = on != 0;
    plan->p->drop_graphs();
    return ICM_OK;
}

int icm_get_graph_status(icm_plan plan, int* status)
{
    if (!plan || !status) return ICM_ERR_INVALID_ARG;
    const PlanBase& P = *plan->p;
    if (P.graph_failed) *status = -1;
    else *status = (P.gexec[2] || (P.gexec[0] && P.gexec[1])) ? 1 : 0;
    return ICM_OK;
}

int icm_get_tuning_left(icm_plan plan, int* forwards)
{
    if (!plan || !forwards) return ICM_ERR_INVALID_ARG;
    *forwards = plan->p->tune_left;
    return ICM_OK;
}

int icm_tuning_candidate(int cand, int* segment_tiles, int* small_tables)
{
    if (cand < -1 || cand >= kNCand) return ICM_ERR_INVALID_ARG;
    if (segment_tiles) *segment_tiles = cand < 0 ? 0 : kSegCand[cand / 2];
    if (small_tables) *small_tables = cand < 0 ? -1 : (cand & 1);
    return ICM_OK;
}

int icm_set_trace(icm_plan plan, int passes)
{
    if (!plan || passes < 0) return ICM_ERR_INVALID_ARG;
    PlanBase& P = *plan->p;
    P.trace_passes = passes;
    P.drop_graphs();   // a captured graph holds the buffers it was captured with
    const int M = P.cfg.num_imfs;
    if (passes > 0) {
        const int64_t nn = (int64_t)(M - 1) * passes * 2 * (int64_t)P.hn.len.size();
        const int64_t ns = (int64_t)(M - 1) * passes * 2 * (int64_t)P.hsig.len.size();
        if (P.alloc(&P.trace[0], nn) || P.alloc(&P.trace[1], ns)) return ICM_ERR_OUT_OF_MEMORY;
        if (P.alloc(&P.trace_hash[0], nn / 2) || P.alloc(&P.trace_hash[1], ns / 2))
            return ICM_ERR_OUT_OF_MEMORY;
        (void)hipMemset(P.trace[0], 0, nn * 4);
        (void)hipMemset(P.trace[1], 0, ns * 4);
        (void)hipMemset(P.trace_hash[0], 0, nn * 4);
        (void)hipMemset(P.trace_hash[1], 0, ns * 4);
    }
    if (P.cfg.stop_mode == ICM_STOP_RILLING && !P.stops[0]) {
        if (P.alloc(&P.stops[0], (int64_t)(M - 1) * P.hn.len.size()) ||
            P.alloc(&P.stops[1], (int64_t)(M - 1) * P.hsig.len.size()))
            return ICM_ERR_OUT_OF_MEMORY;
    }
    return ICM_OK;
}

int icm_get_trace(icm_plan plan, int phase, int32_t* host_out, int64_t count)
{
    if (!plan || !host_out || (phase != 0 && phase != 1)) return ICM_ERR_INVALID_ARG;
    PlanBase& P = *plan->p;
    if (!P.trace[phase]) return ICM_ERR_STATE;
    const int64_t nser = phase == 0 ? (int64_t)P.hn.len.size() : (int64_t)P.hsig.len.size();
    const int64_t n = (int64_t)(P.cfg.num_imfs - 1) * P.trace_passes * 2 * nser;
    if (count < n) return ICM_ERR_INVALID_ARG;
    if (hipMemcpy(host_out, P.trace[phase], n * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return ICM_ERR_HIP;
    return ICM_OK;
}

int icm_get_trace_hash(icm_plan plan, int phase, uint64_t* host_out, int64_t count)
{
    if (!plan || !host_out || (phase != 0 && phase != 1)) return ICM_ERR_INVALID_ARG;
    PlanBase& P = *plan->p;
    if (!P.trace_hash[phase]) return ICM_ERR_STATE;
    const int64_t nser = phase == 0 ? (int64_t)P.hn.len.size() : (int64_t)P.hsig.len.size();
    const int64_t n = (int64_t)(P.cfg.num_imfs - 1) * P.trace_passes * nser;
    if (count < n) return ICM_ERR_INVALID_ARG;
    if (hipMemcpy(host_out, P.trace_hash[phase], n * 8, hipMemcpyDeviceToHost) != hipSuccess)
        return ICM_ERR_HIP;
    return ICM_OK;
}

int icm_get_stops(icm_plan plan, int phase, int32_t* host_out, int64_t count)
{
    if (!plan || !host_out || (phase != 0 && phase != 1)) return ICM_ERR_INVALID_ARG;
    PlanBase& P = *plan->p;
    if (!P.stops[phase]) return ICM_ERR_STATE;
    const int64_t nser = phase == 0 ? (int64_t)P.hn.len.size() : (int64_t)P.hsig.len.size();
    const int64_t n = (int64_t)(P.cfg.num_imfs - 1) * nser;
    if (count < n) return ICM_ERR_INVALID_ARG;
    if (hipMemcpy(host_out, P.stops[phase], n * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return ICM_ERR_HIP;
    return ICM_OK;
}

int icm_get_noise_modes(icm_plan plan, void* d_out, int64_t count, void* stream)
{
    if (!plan || !d_out) return ICM_ERR_INVALID_ARG;
    return plan->p->noise_modes_out(d_out, count, (hipStream_t)stream);
}

int icm_sift_pass(int dtype, const void* d_h, int64_t N, int64_t I, void* d_h_new, int32_t* d_nk,
                  void* stream)
{
    if (!d_h || !d_h_new) return ICM_ERR_INVALID_ARG;
    icm_config c;
    icm_config_default(&c);
    c.dtype = dtype;
    c.num_channels = 1;
    int64_t len = N;
    c.lengths = &len;
    c.num_noise = (int)I;
    c.num_imfs = 2;
    c.max_sifts = 1;
    c.noise_max_sifts = 1;
    c.use_graph = 0;
    if (I < 1 || I > (1 << 20)) return ICM_ERR_INVALID_ARG;
    icm_plan plan = nullptr;
    int rc = icm_plan_create(&c, &plan);
    if (rc) return rc;
    if (dtype == ICM_DTYPE_F64)
        rc = static_cast<Plan<double>*>(plan->p)->sift_once((const double*)d_h, (double*)d_h_new,
                                                            d_nk, (hipStream_t)stream);
    else
        rc = static_cast<Plan<float>*>(plan->p)->sift_once((const float*)d_h, (float*)d_h_new,
                                                           d_nk, (hipStream_t)stream);
    icm_plan_destroy(plan);
    return rc;
}

int icm_noise(int dtype, uint64_t seed, int64_t I, int64_t N, void* d_out, void* stream)
{
    if (!d_out || I < 1 || N < 1 || I > 65535 || N > kMaxLen) return ICM_ERR_INVALID_ARG;
    hipError_t e;
    if (dtype == ICM_DTYPE_F64)
        e = launch_noise<double>(seed, (int)I, N, (double*)d_out, N, (hipStream_t)stream);
    else if (dtype == ICM_DTYPE_F32)
        e = launch_noise<float>(seed, (int)I, N, (float*)d_out, N, (hipStream_t)stream);
    else
        return ICM_ERR_INVALID_ARG;
    if (e != hipSuccess) return ICM_ERR_HIP;
    if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) return ICM_ERR_HIP;
    return ICM_OK;
}

int icm_forward_begin(icm_plan plan, const void* const* d_signals, const void* d_noise,
                      void* stream)
{
    if (!plan || !d_signals) return ICM_ERR_INVALID_ARG;
    return plan->p->forward_begin(d_signals, d_noise, (hipStream_t)stream);
}

int icm_forward_stages(icm_plan plan, int count, int* stage, void* stream)
{
    if (!plan || count < 0) return ICM_ERR_INVALID_ARG;
    return plan->p->forward_stages(count, stage, (hipStream_t)stream);
}

int icm_forward_finish(icm_plan plan, void* const* d_imfs, void* stream)
{
    if (!plan || !d_imfs) return ICM_ERR_INVALID_ARG;
    return plan->p->forward_finish(d_imfs, (hipStream_t)stream);
}

int icm_state_bytes(icm_plan plan, int64_t* bytes)
{
    if (!plan || !bytes) return ICM_ERR_INVALID_ARG;
    *bytes = plan->p->state_bytes();
    return ICM_OK;
}

int icm_state_save(icm_plan plan, void* host, int64_t bytes, void* stream)
{
    if (!plan) return ICM_ERR_INVALID_ARG;
    return plan->p->state_save(host, bytes, (hipStream_t)stream);
}

int icm_state_load(icm_plan plan, const void* host, int64_t bytes, const void* d_noise,
                   void* stream)
{
    if (!plan) return ICM_ERR_INVALID_ARG;
    return plan->p->state_load(host, bytes, d_noise, (hipStream_t)stream);
}

int icm_backward(icm_plan plan, const void* const* d_grad_imfs, void* const* d_grad_signals,
                 void* stream)
{
    if (!plan || !d_grad_imfs || !d_grad_signals) return ICM_ERR_INVALID_ARG;
    return plan->p->backward(d_grad_imfs, d_grad_signals, (hipStream_t)stream);
}

}  // extern "C"
