/*
 * FP64 CPU oracle for Improved CEEMDAN — TEST INFRASTRUCTURE ONLY.
 *
 * This is the checker the HIP path is compared against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product library never links or calls it.
 *
 * It restates the reference algorithm (cudaICEEMDAN.cu:41-617 kernels,
 * :619-1338 host orchestration, statistics.cu:58-131 statistics) with the
 * semantic resolutions listed in iceemdan_oracle.c, and it mirrors the HIP
 * path's arithmetic order exactly (same windowed tridiagonal solve, same
 * canonical reduction order, no FMA contraction) so that FP64 results are
 * bit-identical.
 */
#ifndef ICEEMDAN_ORACLE_H
#define ICEEMDAN_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* stop rules */
#define OCM_STOP_FIXED   0   /* exactly S subtractions, then export */
#define OCM_STOP_RILLING 1   /* reference rule, cap passes (cudaICEEMDAN.cu:494-544) */

/* spline boundary condition (process-wide setting of the oracle) */
#define OCM_BC_NATURAL     0
#define OCM_BC_NOT_A_KNOT  1
void ocm_set_boundary(int bc);

/* Strict local extrema with both endpoints forced (cudaICEEMDAN.cu:92-148). */
int ocm_extrema(const double* h, int N, int is_max, int* pos, double* val);

/* Endpoint extrapolation fix when n >= 6 (cudaICEEMDAN.cu:209-291). */
void ocm_boundary_fix(int n, const int* pos, double* val, int is_max);

/* Natural-spline second-derivative/2 coefficients c[0..n) by the windowed
 * partition + PCR solve (64 lanes x B rows per window, H halo rows). */
void ocm_spline_solve(int n, const int* pos, const double* val, int B, int H, double* c);

/* Same system solved globally by the Thomas algorithm (validation only). */
void ocm_spline_solve_thomas(int n, const int* pos, const double* val, double* c);

/* One window of the partition + PCR solve: rows g0 .. g0+64B-1, all W values. */
void ocm_solve_window(int n, const int* pos, const double* val, int B, long g0, double* x);

/* Envelope at every sample 0..N-1 (cudaICEEMDAN.cu:382-439). */
void ocm_envelope(int N, int n, const int* pos, const double* val, const double* c, double* env);

/* One sift pass.  Returns 0.  h_new = h - m.  nmax/nmin: knot counts.
 * cnt_sx / any_mark: Rilling statistics of this pass (may be NULL). */
int ocm_sift_pass(const double* h, int N, int B, int H, double* h_new,
                  int* nmax, int* nmin, int64_t* cnt_sx, int* any_mark);

/* Rilling continue decision for one realization (cudaICEEMDAN.cu:494-525). */
int ocm_rilling_continue(int64_t cnt_sx, int any_mark, int nmax, int nmin, int N);

/* Canonical reductions (1024 strided lanes, then a pairwise tree). */
double ocm_sum_canon(const double* x, long N);
double ocm_var_canon(const double* x, long N);   /* unbiased, two-pass */

/* Extract one mode from x (sifting). Writes the exported mode, returns the
 * number of subtractions performed.  trace (optional): per pass nmax,nmin. */
int ocm_sift_mode(const double* x, int N, int stop_mode, int S_or_cap, int B, int H,
                  double* mode, double* work /*3N*/, int* trace /*2*passes*/, int trace_passes);

/* Full ICEEMDAN.  noise: I x N standard normals (realization-major).
 * imfs: M x N.  trace (optional): int[2][M-1][trace_passes][I][2] knot counts,
 * stops (optional): int[2][M-1][I] subtractions performed.
 * noise_modes_out (optional): (M-1) x I x N noise modes.
 * Returns 0 on success, <0 on bad arguments. */
int ocm_iceemdan(const double* y, int N, int I, int M, int stop_mode, int S_or_cap,
                 int noise_cap, double eps0, const double* noise, int B, int H, int nthreads,
                 double* imfs, int* trace, int trace_passes, int* stops, double* noise_modes_out);

/* Adjoint of one envelope operator with frozen knots: given dL/d(envelope)
 * g[N], accumulates into out[N] (+=) the gradient with respect to the
 * sifted samples h.  pos/n: knot positions; fixed_lo/fixed_hi: whether the
 * endpoint fix replaced the value of knot 0 / knot n-1 in the forward. */
void ocm_envelope_adjoint(int N, int n, const int* pos, int fixed_lo, int fixed_hi,
                          const double* g, double* out);

/* Gradient of L(IMFs) with respect to the input channel y, FIXED stop rule,
 * extrema sets and endpoint-fix branches frozen at the forward's values.
 * grad_imfs: dL/dIMF, M x N.  grad_y: N (output).  Also returns the IMFs. */
int ocm_iceemdan_grad(const double* y, int N, int I, int M, int S, double eps0,
                      const double* noise, int B, int H, int nthreads, const double* grad_imfs,
                      double* grad_y, double* imfs);

/* Philox4x64-10 block (Salmon et al. SC'11) and the Gaussian stream built on it. */
void ocm_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);
void ocm_noise(uint64_t seed, long I, long N, double* out);

#ifdef __cplusplus
}
#endif
#endif
