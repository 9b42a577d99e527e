/*
 * C ABI of the MI355X-native Improved CEEMDAN library (libiceemdan_mi355x.so).
 *
 * All data pointers named d_* are device pointers owned by the caller.
 * Streams are passed as `void*` (a hipStream_t; NULL = the null stream).
 * Every entry point returns an ICM_* status code (0 = success).
 *
 * Reference interfaces replaced (EEGLab-Pannon/CEEMDAN-GPU):
 *   icm_plan_create/icm_forward/icm_plan_destroy
 *       one call of iceemdanS / iceemdanD (cudaICEEMDAN.cuh:25-26, defined
 *       cudaICEEMDAN.cu:619 and :1340), generalised to many channels of
 *       uneven length and to a reusable plan (allocation hoisted out of the
 *       call, cudaICEEMDAN.cu:621-765).
 *   icm_noise         curandGenerateNormal(Double) (cudaICEEMDAN.cu:739-741,
 *                     :1460-1462), replaced by Philox4x64-10 + Box-Muller.
 *   icm_sift_pass     one iteration of the sifting loop body
 *                     (cudaICEEMDAN.cu:1031-1137) for I series of length N.
 *   icm_backward      no reference counterpart (adjoint of the decomposition).
 * The drop-in C++ entry points iceemdanS/iceemdanD are declared in
 * include/cudaICEEMDAN.h and the statistics kernels in include/statistics.h.
 */
#ifndef ICEEMDAN_MI355X_H
#define ICEEMDAN_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ICM_OK 0
#define ICM_ERR_INVALID_ARG (-1)
#define ICM_ERR_OUT_OF_MEMORY (-2)
#define ICM_ERR_HIP (-3)
#define ICM_ERR_DEVICE (-4)      /* a kernel raised its device error word */
#define ICM_ERR_INDEX (-5)       /* sample-index array is not 0..N-1 */
#define ICM_ERR_STATE (-6)       /* call not valid in the plan's state */

#define ICM_DTYPE_F32 0
#define ICM_DTYPE_F64 1

#define ICM_STOP_FIXED 0         /* exactly max_sifts subtractions per stage */
#define ICM_STOP_RILLING 1       /* reference stop rule, capped at max_sifts */

#define ICM_BOUNDARY_NATURAL 0      /* natural spline ends (the reference's) */
#define ICM_BOUNDARY_NOT_A_KNOT 1   /* not-a-knot ends (cudaICEEMDAN.cu:345-380 was never wired) */

#define ICM_PHASE_NOISE 0
#define ICM_PHASE_SIGNAL 1

typedef struct icm_config {
    int dtype;                  /* ICM_DTYPE_F32 / ICM_DTYPE_F64 */
    int num_channels;           /* 1 <= C <= 65535 */
    const int64_t* lengths;     /* host array [C]: samples per channel (4 .. 16,000,000) */
    int num_noise;              /* I >= 1: noise realizations per channel */
    int num_imfs;               /* M >= 2: output rows (M-1 sifted stages + residue) */
    int stop_mode;              /* ICM_STOP_FIXED / ICM_STOP_RILLING */
    int max_sifts;              /* FIXED: sifts per stage; RILLING: cap (reference max_iter);
                                   1 .. 32000 */
    int noise_max_sifts;        /* the same for the noise-mode stages (reference: 2000), 1 .. 32000 */
    double noise_strength;      /* eps0 */
    uint64_t seed;              /* Philox key of the white noise */
    int use_graph;              /* replay the forward as a captured hipGraph (FIXED; RILLING with
                                   both caps <= 64 passes) */
    int boundary;               /* ICM_BOUNDARY_*: spline end condition (envelopes with >= 4
                                   knots; fewer knots keep natural ends) */
    int keep_records;           /* FIXED: the forward keeps every pass's knot structure so that
                                   icm_backward need not re-sift (memory: about
                                   (M-1) * max_sifts * 4 * I * sum(N) bytes) */
} icm_config;

typedef struct icm_plan_s* icm_plan;

typedef struct icm_plan_info {
    int64_t device_bytes;       /* bytes allocated by the plan */
    int64_t noise_series;       /* noise-phase series (noise groups x I) */
    int64_t signal_series;      /* signal-phase series (C x I) */
    int64_t noise_samples;      /* samples per noise-phase pass */
    int64_t signal_samples;     /* samples per signal-phase pass */
    int64_t tiles_noise;
    int64_t tiles_signal;
    int32_t tile;               /* samples per tile */
    int32_t solve_rows_per_lane;/* B of the windowed solve */
    int32_t solve_halo;         /* H of the windowed solve */
} icm_plan_info;

typedef struct icm_timing {
    double noise_ms;            /* noise generation + noise-mode stages (-1: not measured) */
    double signal_ms;           /* CEEMDAN stages (reference timer, cudaICEEMDAN.cu:1002-1297;
                                   with the phases side by side it includes waits for the
                                   noise modes; -1: not measured) */
    double total_ms;            /* whole forward, input copy to output copy */
    double backward_ms;         /* last icm_backward, gradient in to gradient out */
} icm_timing;

const char* icm_status_string(int status);
/* Detail of the last HIP runtime failure on the calling thread ("" if none). */
const char* icm_last_error(void);

/* Host-only checks (no device calls). */
int icm_config_default(icm_config* cfg);
int icm_validate_config(const icm_config* cfg);
int icm_plan_bytes(const icm_config* cfg, int64_t* bytes);

/* Plan lifecycle.  A plan owns all scratch; it is reentrant across plans but
 * a single plan must not run two forwards concurrently. */
int icm_plan_create(const icm_config* cfg, icm_plan* out);
int icm_plan_destroy(icm_plan plan);
int icm_plan_get_info(icm_plan plan, icm_plan_info* info);

/* Forward decomposition.
 * d_signals[c]: channel c, lengths[c] samples of dtype.
 * d_imfs[c]:    channel c output, [num_imfs][lengths[c]] row-major; rows
 *               0..M-2 are the modes, row M-1 the residue, all in units of
 *               x / std(x) (cudaICEEMDAN.cu:1013-1020, :1298).
 * d_noise:      optional [I][N] standard normals (all channels the same
 *               length N); NULL = generate from cfg.seed.
 * Synchronous: returns after the result is complete. */
int icm_forward(icm_plan plan, const void* const* d_signals, void* const* d_imfs,
                const void* d_noise, void* stream);
int icm_last_timing(icm_plan plan, icm_timing* out);

/* FIXED mode runs the noise-mode phase on a second stream of the plan beside
 * the CEEMDAN phase (stage k of which waits only for noise-mode stage k), so
 * that one phase's latency-bound launches run beside the other's sift
 * passes; the result is the same bits.  on = 0 runs the phases one after
 * the other (also ICM_OVERLAP=0 in the environment when a plan is made).
 * Tuning and traced forwards always run one phase after the other.  With
 * the phases side by side and a captured graph, icm_last_timing reports
 * noise_ms = signal_ms = -1 (only the total is measured). */
int icm_set_overlap(icm_plan plan, int on);

/* Graph state of the plan after its last forward: 1 = both phases replay a
 * captured hipGraph; 0 = launched eagerly (use_graph off, or the forward was
 * tuning, profiled or traced, or RILLING caps above 64 passes); -1 = a
 * capture failed and the plan runs eagerly from then on (the reason is in
 * icm_last_error()). */
int icm_get_graph_status(icm_plan plan, int* status);

/* Per-kernel timing with HIP events on the forward's stream (disables the
 * hipGraph replay while on).  kind: 0 = stage init (extrema of the stage
 * input), 1 = spline solve, 2 = sift pass.  Totals of the last forward;
 * alg_bytes = algorithmic bytes (2 * sizeof(real) per sample per launch). */
#define ICM_KERNEL_INIT 0
#define ICM_KERNEL_SOLVE 1
#define ICM_KERNEL_SIFT 2
int icm_set_profiling(icm_plan plan, int on);
int icm_get_kernel_stats(icm_plan plan, int kind, double* total_ms, int64_t* launches,
                         int64_t* alg_bytes);

/* Autotuner of the sift pass (no reference counterpart: it replaces the
 * reference's fixed launch geometry, cudaICEEMDAN.cu:767-830).  For each
 * phase (noise modes, CEEMDAN) and stage it picks the sift's segment length
 * (consecutive tiles per work-queue item: 4, 8, 16 or 32) and knot-table size
 * (full tables, or small tables with a second full-table launch for the
 * tiles that overflow them) by timing the candidates, in rotation over the
 * passes, during the plan's next `forwards` icm_forward calls (HIP events on
 * the forward's stream; those forwards run without the hipGraph, which is
 * captured once tuning ends).  The choice is kept in a process-wide table
 * keyed by the plan's shape and device, so later plans of the same shape
 * start tuned.  Every candidate computes the same bits.  forwards = 0: the
 * default geometry (segments of clamp(tiles / 56000, 1, 32) tiles, small
 * tables from the second stage on).  New plans tune over 2 forwards, or over
 * ICM_AUTOTUNE forwards when that environment variable is set (0: off). */
int icm_set_autotune(icm_plan plan, int forwards);
/* Forwards that will still time candidates (0: tuned or tuning off). */
int icm_get_tuning_left(icm_plan plan, int* forwards);
/* Candidate per stage of a phase (host_out[k], k < M-1; -1 = default). */
int icm_get_tuning(icm_plan plan, int phase, int32_t* host_out, int64_t count);
/* Sets every stage's candidate (-1 .. 7; count = M-1) and ends tuning. */
int icm_set_tuning(icm_plan plan, int phase, const int32_t* host_in, int64_t count);
/* Meaning of candidate `cand`: tiles per segment (0 = default) and small
 * knot tables (1), full tables (0) or the default rule (-1). */
int icm_tuning_candidate(int cand, int* segment_tiles, int* small_tables);

/* Per-pass knot counts (debug/parity): enable before a forward with the
 * number of passes to keep per stage; read back [M-1][passes][series][2]
 * (n_max, n_min of each pass input) of a phase. */
int icm_set_trace(icm_plan plan, int passes);
int icm_get_trace(icm_plan plan, int phase, int32_t* host_out, int64_t count);
/* Knot-set digests of the same traced passes: [M-1][passes][series] 64-bit
 * order-independent hashes of both knot sets of each pass input (equal
 * digests: the same extrema index sets, up to a hash collision). */
int icm_get_trace_hash(icm_plan plan, int phase, uint64_t* host_out, int64_t count);
/* RILLING: subtractions performed per [M-1][series] of a phase. */
int icm_get_stops(icm_plan plan, int phase, int32_t* host_out, int64_t count);
/* noise modes E_1..E_{M-1} of the last forward: [M-1][noise series][N] */
int icm_get_noise_modes(icm_plan plan, void* d_out, int64_t count, void* stream);

/* Standard-normal generator: out[i][p], i < I, p < N (Philox4x64-10, key
 * (seed, 0), counter (p/4, i, 0, 0), Box-Muller). */
int icm_noise(int dtype, uint64_t seed, int64_t I, int64_t N, void* d_out, void* stream);

/* One sift pass over I series of length N stored back to back:
 * d_h_new = d_h - mean envelope; d_nk[2*i] / d_nk[2*i+1] = knot counts
 * (maxima, minima) of series i of d_h (may be NULL). */
int icm_sift_pass(int dtype, const void* d_h, int64_t N, int64_t I, void* d_h_new, int32_t* d_nk,
                  void* stream);

/* icm_sift_pass without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, kernels and result bit
 * for bit as icm_sift_pass; every buffer beyond d_h, d_h_new and d_nk is
 * carved from one file-scope device array of 24 MiB per GPU, held until the
 * process ends.  Limits: I * N <= 2^21 samples and I <= 4096 series, else
 * ICM_ERR_INVALID_ARG.  Calls on one GPU share that GPU's array (it keeps
 * the last call's contents), so calls that overlap on one GPU race: the
 * caller serialises them.  Synchronous. */
int icm_sift_pass_static(int dtype, const void* d_h, int64_t N, int64_t I, void* d_h_new,
                         int32_t* d_nk, void* stream);

/* Staged forward with stage-boundary checkpoints (no reference counterpart;
 * the reference runs a decomposition in one call and cannot resume one).
 *   icm_forward_begin   inputs, noise modes and stage-0 statistics; the plan
 *                       is then before CEEMDAN stage 0
 *   icm_forward_stages  runs up to `count` CEEMDAN stages; *stage (may be
 *                       NULL) receives the next stage, M-1 when all are done
 *   icm_forward_finish  after the last stage: residue row, copy to d_imfs
 * The result equals icm_forward's bit for bit.  Between stages the state
 * (stage k, the inputs, r_k, rows 0..k-1, the residue history) can be
 * copied to host memory and loaded into a plan of the same configuration,
 * in this or another process; loading recomputes the noise modes from the
 * seed, or from d_noise when the saved run was given its noise (pass the
 * same noise again).  icm_state_bytes gives the size of the host buffer. */
int icm_forward_begin(icm_plan plan, const void* const* d_signals, const void* d_noise,
                      void* stream);
int icm_forward_stages(icm_plan plan, int count, int* stage, void* stream);
int icm_forward_finish(icm_plan plan, void* const* d_imfs, void* stream);
int icm_state_bytes(icm_plan plan, int64_t* bytes);
int icm_state_save(icm_plan plan, void* host, int64_t bytes, void* stream);
int icm_state_load(icm_plan plan, const void* host, int64_t bytes, const void* d_noise,
                   void* stream);

/* Adjoint of the last icm_forward of this plan (no reference counterpart):
 * given dL/dIMF (d_grad_imfs[c]: [M][N_c], the layout of d_imfs), writes
 * dL/dx (d_grad_signals[c]: [N_c]) for every channel.  FIXED stop rule only
 * (ICM_ERR_STATE otherwise, or when no forward ran).  The extrema sets and
 * endpoint-fix branches are those of the forward: each stage's sifting is
 * re-run from the saved residue to record them, then the passes are
 * reversed.  The noise is treated as a constant.  Device memory for the
 * records is allocated on the first call and kept by the plan. */
int icm_backward(icm_plan plan, const void* const* d_grad_imfs, void* const* d_grad_signals,
                 void* stream);

#ifdef __cplusplus
}
#endif
#endif
