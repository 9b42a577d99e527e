/*
 * FP64 CPU oracle for Improved CEEMDAN — TEST INFRASTRUCTURE ONLY.
 *
 * Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
 * the checker.  The product library (ceemdan-gpu_amd/) never links it.
 *
 * Build: gcc -O2 -ffp-contract=off -mfma -fopenmp (see oracle/build_oracle.py).
 * No FMA contraction and no fast-math anywhere: every expression below is
 * evaluated in IEEE double, in the written order, exactly as the HIP kernels
 * evaluate it, so FP64 results are bit-identical to the GPU path.
 *
 * What each function restates (reference file:line) and how it is pinned
 * (tests/test_oracle.py):
 *   ocm_extrema         cudaICEEMDAN.cu:92-148 (strict >/<, endpoints forced)
 *                         pinned: brute force on random/plateau/NaN inputs
 *   ocm_boundary_fix    cudaICEEMDAN.cu:209-291 (n >= 6, linear extrapolation)
 *                         pinned: hand-computed cases
 *   ocm_spline_solve    tridiagonal_setup :310-343 + gtsv2 solve :899-921
 *                         pinned: scipy CubicSpline(bc_type='natural'),
 *                         scipy.linalg.solve_banded, global Thomas, residual
 *   ocm_envelope        spline_coefficients :382-401 + interpolate :403-439
 *                         pinned: scipy CubicSpline evaluation, knot interpolation,
 *                         exactness on linear data
 *   ocm_sift_pass       averageUppperLower :441-454, produceSX :456-467,
 *                       thresholdJudge :469-492, updateRealizations :527-544
 *                         pinned: composed from the pinned parts; hand cases
 *   ocm_rilling_continue siftingCriterion :494-525 (C precedence kept)
 *                         pinned: hand-built truth table
 *   ocm_sum/var_canon   statistics.cu:58-131 (mean / variance, unbiased)
 *                         pinned: numpy mean/var(ddof=1) to round-off
 *   ocm_iceemdan        host drivers :845-994 (noise modes), :1003-1298 (stages)
 *                         pinned: completeness sum(rows) == x/std(x),
 *                         eps0 = 0 independence of I and noise, golden fixtures
 *   ocm_philox4x64      Random123 Philox4x64-10 (replaces cuRAND XORWOW,
 *                       cudaICEEMDAN.cu:739-741): pinned against numpy.random.Philox
 *                       and the published known-answer vectors.
 *
 * Semantic resolutions of undefined reference behaviour (same in HIP path):
 *   - mean envelope when n_max <= 3 or n_min <= 3: the reference leaves the
 *     buffer stale/uninitialised (:450); here m = 0 (no subtraction).
 *   - Rilling fraction: count(sx > 0.05) * (1/N) per pass; the reference's
 *     accumulator is never reset (:975-976).
 *   - stage statistics: exact two-pass unbiased mean/variance per stage; the
 *     reference's accumulators are never reset (:1011-1021, :1153-1157).
 *   - reductions in a fixed canonical order instead of float atomics.
 *   - early exit: as if every pass up to the cap ran with per-realization masks.
 *   - tridiagonal solve: windows of W = 64*B rows keep their central
 *     K = W - 2H rows; coupling beyond a window edge is dropped.  The
 *     off-diagonal/diagonal ratio of the natural-spline system is exactly 1/2,
 *     so the truncation error at H rows from an edge is below 2^(1-H) of the
 *     solution scale (H = 64 for FP64): beneath double round-off.
 */
#include "iceemdan_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NLANES 64

/* ------------------------------------------------------------------ extrema */

int ocm_extrema(const double* h, int N, int is_max, int* pos, double* val)
{
    int n = 0;
    for (int p = 0; p < N; ++p) {
        int flag;
        if (p == 0 || p == N - 1) {
            flag = 1;
        } else if (is_max) {
            flag = (h[p] > h[p - 1]) && (h[p] > h[p + 1]);
        } else {
            flag = (h[p] < h[p - 1]) && (h[p] < h[p + 1]);
        }
        if (flag) {
            pos[n] = p;
            val[n] = h[p];
            ++n;
        }
    }
    return n;
}

/* -------------------------------------------------------- boundary fix-up */

void ocm_boundary_fix(int n, const int* pos, double* val, int is_max)
{
    if (n < 6) return;
    /* left end: extrapolate the line through knots 1,2 back to knot 0 */
    double slope = (val[2] - val[1]) / (double)(pos[2] - pos[1]);
    double t = val[1] - slope * (double)(pos[1] - pos[0]);
    /* right end: extrapolate the line through knots n-3,n-2 to knot n-1 */
    double slope_r = (val[n - 2] - val[n - 3]) / (double)(pos[n - 2] - pos[n - 3]);
    double t_r = val[n - 2] + slope_r * (double)(pos[n - 1] - pos[n - 2]);
    if (is_max) {
        if (t > val[0]) val[0] = t;
        if (t_r > val[n - 1]) val[n - 1] = t_r;
    } else {
        if (t < val[0]) val[0] = t;
        if (t_r < val[n - 1]) val[n - 1] = t_r;
    }
}

/* ----------------------------------------------------- tridiagonal system */

/* Boundary condition of the spline (natural by default; not-a-knot is the
 * option the reference started but never wired, cudaICEEMDAN.cu:345-380). */
static int g_bc = OCM_BC_NATURAL;
void ocm_set_boundary(int bc) { g_bc = bc; }

/* Row g of the spline system on knots (pos, val), n rows.  Rows outside
 * [0, n) and the two end rows are inert identity rows (the reference pads
 * with them, cudaICEEMDAN.cu:293-308).  Not-a-knot (n >= 4): d_0 = d_1 and
 * d_{n-3} = d_{n-2}; c_0 is eliminated from row 1 and c_{n-1} from row n-2,
 * which keeps both rows diagonally dominant, and the end values are
 * recovered after the solve (ocm_nak_end). */
static void spline_row(int n, const int* pos, const double* y, long g,
                       double* dl, double* d, double* du, double* f)
{
    if (g < 0 || g >= n || g == 0 || g == n - 1) {
        *dl = 0.0; *d = 1.0; *du = 0.0; *f = 0.0;
        return;
    }
    int hl = pos[g] - pos[g - 1];
    int hr = pos[g + 1] - pos[g];
    *dl = (double)hl;
    *du = (double)hr;
    *d = (double)(2 * (hl + hr));
    *f = 3.0 * ((y[g + 1] - y[g]) / (double)hr - (y[g] - y[g - 1]) / (double)hl);
    if (g_bc == OCM_BC_NOT_A_KNOT && n >= 4) {
        if (g == 1) {
            *dl = 0.0;
            *d = (double)(3 * hl + 2 * hr) + (double)hl * (double)hl / (double)hr;
            *du = (double)hr - (double)hl * (double)hl / (double)hr;
        }
        if (g == n - 2) {
            *dl = (double)hl - (double)hr * (double)hr / (double)hl;
            *d = (double)(2 * hl + 3 * hr) + (double)hr * (double)hr / (double)hl;
            *du = 0.0;
        }
    }
}

/* Not-a-knot end value from its two neighbours: c_e = c_1 (1 + r) - c_2 r,
 * r = h_e / h_next (h_e: interval at the end, h_next: the one after it). */
static double nak_end(double c1, double c2, int h_e, int h_next)
{
    double r = (double)h_e / (double)h_next;
    return c1 * (1.0 + r) - c2 * r;
}

void ocm_solve_window(int n, const int* pos, const double* y, int B, long g0, double* x)
{
    /* Per-lane interior elimination: rows q = 0..B-2 of lane l expressed as
     * x_q = alpha_q + beta_q * z_{l-1} + gamma_q * z_l, z_l = row B-1 of lane l. */
    double* al = (double*)malloc(sizeof(double) * NLANES * B * 3);
    double* be = al + NLANES * B;
    double* ga = be + NLANES * B;
    double A[NLANES], D[NLANES], C[NLANES], R[NLANES];
    double cp[64];
    double aa[64], bb[64], gg[64];

    for (int l = 0; l < NLANES; ++l) {
        long rb = g0 + (long)l * B;
        double dl, d, du, f;
        /* forward sweep over interior rows */
        for (int q = 0; q <= B - 2; ++q) {
            spline_row(n, pos, y, rb + q, &dl, &d, &du, &f);
            double inv;
            if (q == 0) {
                /* lane 0's first row couples outside the window: dropped */
                double dl0 = (l == 0) ? 0.0 : dl;
                inv = 1.0 / d;
                aa[q] = f * inv;
                bb[q] = (-dl0) * inv;
            } else {
                inv = 1.0 / (d - dl * cp[q - 1]);
                aa[q] = (f - dl * aa[q - 1]) * inv;
                bb[q] = (-dl) * bb[q - 1] * inv;
            }
            cp[q] = du * inv;
            /* the last interior row couples to z_l through its du */
            gg[q] = (q == B - 2) ? (-du) * inv : 0.0;
        }
        /* back substitution */
        double* a_l = al + (long)l * B;
        double* b_l = be + (long)l * B;
        double* g_l = ga + (long)l * B;
        a_l[B - 2] = aa[B - 2];
        b_l[B - 2] = bb[B - 2];
        g_l[B - 2] = gg[B - 2];
        for (int q = B - 3; q >= 0; --q) {
            a_l[q] = aa[q] - cp[q] * a_l[q + 1];
            b_l[q] = bb[q] - cp[q] * b_l[q + 1];
            g_l[q] = gg[q] - cp[q] * g_l[q + 1];
        }
    }
    /* interface rows -> reduced tridiagonal system in z_0..z_63 */
    for (int l = 0; l < NLANES; ++l) {
        long rb = g0 + (long)l * B;
        double dl, d, du, f;
        spline_row(n, pos, y, rb + B - 1, &dl, &d, &du, &f);
        double na0 = 0.0, nb0 = 0.0, ng0 = 0.0;
        if (l == NLANES - 1) {
            du = 0.0;
        } else {
            na0 = al[(long)(l + 1) * B];
            nb0 = be[(long)(l + 1) * B];
            ng0 = ga[(long)(l + 1) * B];
        }
        double a_e = al[(long)l * B + B - 2];
        double b_e = be[(long)l * B + B - 2];
        double g_e = ga[(long)l * B + B - 2];
        A[l] = dl * b_e;
        D[l] = (d + dl * g_e) + du * nb0;
        C[l] = du * ng0;
        R[l] = (f - dl * a_e) - du * na0;
    }
    /* parallel cyclic reduction, strides 1..32 */
    for (int s = 1; s < NLANES; s <<= 1) {
        double A2[NLANES], D2[NLANES], C2[NLANES], R2[NLANES];
        for (int l = 0; l < NLANES; ++l) {
            int lo = l - s, hi = l + s;
            double a2 = 0.0, c2 = 0.0, d2 = D[l], r2 = R[l];
            if (lo >= 0) {
                double k1 = A[l] / D[lo];
                a2 = -(A[lo] * k1);
                d2 = d2 - C[lo] * k1;
                r2 = r2 - R[lo] * k1;
            }
            if (hi < NLANES) {
                double k2 = C[l] / D[hi];
                c2 = -(C[hi] * k2);
                d2 = d2 - A[hi] * k2;
                r2 = r2 - R[hi] * k2;
            }
            A2[l] = a2; C2[l] = c2; D2[l] = d2; R2[l] = r2;
        }
        memcpy(A, A2, sizeof A); memcpy(C, C2, sizeof C);
        memcpy(D, D2, sizeof D); memcpy(R, R2, sizeof R);
    }
    double z[NLANES];
    for (int l = 0; l < NLANES; ++l) z[l] = R[l] / D[l];
    for (int l = 0; l < NLANES; ++l) {
        double zl = (l > 0) ? z[l - 1] : 0.0;
        for (int q = 0; q <= B - 2; ++q) {
            long k = (long)l * B + q;
            x[k] = (al[k] + be[k] * zl) + ga[k] * z[l];
        }
        x[(long)l * B + B - 1] = z[l];
    }
    free(al);
}

void ocm_spline_solve(int n, const int* pos, const double* val, int B, int H, double* c)
{
    long W = (long)NLANES * B;
    long K = W - 2L * H;
    double* x = (double*)malloc(sizeof(double) * W);
    for (long w = 0; w * K < n; ++w) {
        long g0 = w * K - H;
        ocm_solve_window(n, pos, val, B, g0, x);
        for (long r = H; r < H + K; ++r) {
            long g = g0 + r;
            if (g >= 0 && g < n) c[g] = x[r];
        }
        if (g_bc == OCM_BC_NOT_A_KNOT && n >= 4) {
            /* the window that keeps an end knot recovers it from its own rows */
            long r0 = -g0, r1 = (long)(n - 1) - g0;
            if (r0 >= H && r0 < H + K)
                c[0] = nak_end(x[r0 + 1], x[r0 + 2], pos[1] - pos[0], pos[2] - pos[1]);
            if (r1 >= H && r1 < H + K)
                c[n - 1] = nak_end(x[r1 - 1], x[r1 - 2], pos[n - 1] - pos[n - 2],
                                   pos[n - 2] - pos[n - 3]);
        }
    }
    free(x);
}

void ocm_spline_solve_thomas(int n, const int* pos, const double* val, double* c)
{
    double* cp = (double*)malloc(sizeof(double) * n * 2);
    double* dp = cp + n;
    for (int g = 0; g < n; ++g) {
        double dl, d, du, f;
        spline_row(n, pos, val, g, &dl, &d, &du, &f);
        if (g == 0) {
            cp[g] = du / d;
            dp[g] = f / d;
        } else {
            double den = d - dl * cp[g - 1];
            cp[g] = du / den;
            dp[g] = (f - dl * dp[g - 1]) / den;
        }
    }
    c[n - 1] = dp[n - 1];
    for (int g = n - 2; g >= 0; --g) c[g] = dp[g] - cp[g] * c[g + 1];
    if (g_bc == OCM_BC_NOT_A_KNOT && n >= 4) {
        c[0] = nak_end(c[1], c[2], pos[1] - pos[0], pos[2] - pos[1]);
        c[n - 1] = nak_end(c[n - 2], c[n - 3], pos[n - 1] - pos[n - 2], pos[n - 2] - pos[n - 3]);
    }
    free(cp);
}

/* ----------------------------------------------------------- evaluation */

/* Interval coefficients of knot interval k (cudaICEEMDAN.cu:382-401). */
static void interval_coeffs(const int* pos, const double* a, const double* c, int k,
                            double* b, double* d)
{
    double hk = (double)(pos[k + 1] - pos[k]);
    *b = (a[k + 1] - a[k]) / hk - hk * (2.0 * c[k] + c[k + 1]) / 3.0;
    *d = (c[k + 1] - c[k]) / (3.0 * hk);
}

void ocm_envelope(int N, int n, const int* pos, const double* val, const double* c, double* env)
{
    /* Sample p lies in (pos[k], pos[k+1]]; p = 0 uses k = 0 (:418-433). */
    int k = 0;
    double b, d;
    interval_coeffs(pos, val, c, 0, &b, &d);
    for (int p = 0; p < N; ++p) {
        while (k + 1 < n - 1 && p > pos[k + 1]) {
            ++k;
            interval_coeffs(pos, val, c, k, &b, &d);
        }
        double t = (double)(p - pos[k]);
        /* fused multiply-adds, as nvcc contracts the reference's expression
           (:436) and as the HIP kernels evaluate it */
        env[p] = fma(fma(fma(d, t, c[k]), t, b), t, val[k]);
    }
}

/* ------------------------------------------------------------- sift pass */

int ocm_rilling_continue(int64_t cnt_sx, int any_mark, int nmax, int nmin, int N)
{
    double frac = (double)cnt_sx * (1.0 / (double)N);
    return (frac > 0.05) || (any_mark && (nmax + nmin) > 6);
}

int ocm_sift_pass(const double* h, int N, int B, int H, double* h_new,
                  int* nmax_out, int* nmin_out, int64_t* cnt_sx, int* any_mark)
{
    int* pos = (int*)malloc(sizeof(int) * N);
    double* val = (double*)malloc(sizeof(double) * N * 4);
    double* c = val + N;
    double* U = c + N;
    double* L = U + N;

    int nmax = ocm_extrema(h, N, 1, pos, val);
    ocm_boundary_fix(nmax, pos, val, 1);
    ocm_spline_solve(nmax, pos, val, B, H, c);
    ocm_envelope(N, nmax, pos, val, c, U);

    int nmin = ocm_extrema(h, N, 0, pos, val);
    ocm_boundary_fix(nmin, pos, val, 0);
    ocm_spline_solve(nmin, pos, val, B, H, c);
    ocm_envelope(N, nmin, pos, val, c, L);

    int use_mean = (nmax > 3) && (nmin > 3);
    int64_t cnt = 0;
    int mark = 0;
    for (int p = 0; p < N; ++p) {
        double m = use_mean ? (U[p] + L[p]) / 2.0 : 0.0;
        h_new[p] = h[p] - m;
        double sx = fabs((U[p] + L[p]) / (U[p] - L[p]));
        if (sx > 0.5) mark = 1;
        if (sx > 0.05) ++cnt;
    }
    if (nmax_out) *nmax_out = nmax;
    if (nmin_out) *nmin_out = nmin;
    if (cnt_sx) *cnt_sx = cnt;
    if (any_mark) *any_mark = mark;
    free(pos);
    free(val);
    return 0;
}

/* ----------------------------------------------------------- reductions */

#define CANON_LANES 1024

double ocm_sum_canon(const double* x, long N)
{
    double acc[CANON_LANES];
    for (int l = 0; l < CANON_LANES; ++l) acc[l] = 0.0;
    for (long p = 0; p < N; ++p) acc[p % CANON_LANES] = acc[p % CANON_LANES] + x[p];
    for (int s = CANON_LANES / 2; s >= 1; s >>= 1)
        for (int l = 0; l < s; ++l) acc[l] = acc[l] + acc[l + s];
    return acc[0];
}

double ocm_var_canon(const double* x, long N)
{
    double mu = ocm_sum_canon(x, N) * (1.0 / (double)N);
    double acc[CANON_LANES];
    for (int l = 0; l < CANON_LANES; ++l) acc[l] = 0.0;
    for (long p = 0; p < N; ++p) {
        double dv = x[p] - mu;
        acc[p % CANON_LANES] = acc[p % CANON_LANES] + dv * dv;
    }
    for (int s = CANON_LANES / 2; s >= 1; s >>= 1)
        for (int l = 0; l < s; ++l) acc[l] = acc[l] + acc[l + s];
    return acc[0] * (1.0 / (double)(N - 1));
}

/* ---------------------------------------------------------- mode extraction */

int ocm_sift_mode(const double* x, int N, int stop_mode, int S_or_cap, int B, int H,
                  double* mode, double* work, int* trace, int trace_passes)
{
    double* h = work;
    double* hn = work + N;
    memcpy(h, x, sizeof(double) * N);
    int subs = 0;
    if (stop_mode == OCM_STOP_FIXED) {
        for (int j = 0; j < S_or_cap; ++j) {
            int nmax, nmin;
            ocm_sift_pass(h, N, B, H, hn, &nmax, &nmin, NULL, NULL);
            if (trace && j < trace_passes) { trace[2 * j] = nmax; trace[2 * j + 1] = nmin; }
            double* t = h; h = hn; hn = t;
            ++subs;
        }
    } else {
        for (int j = 0; j < S_or_cap; ++j) {
            int nmax, nmin, mark;
            int64_t cnt;
            ocm_sift_pass(h, N, B, H, hn, &nmax, &nmin, &cnt, &mark);
            if (trace && j < trace_passes) { trace[2 * j] = nmax; trace[2 * j + 1] = nmin; }
            int cont = ocm_rilling_continue(cnt, mark, nmax, nmin, N);
            if (!cont || j == S_or_cap - 1) break;
            double* t = h; h = hn; hn = t;
            ++subs;
        }
    }
    memcpy(mode, h, sizeof(double) * N);
    return subs;
}

/* ------------------------------------------------------------ full ICEEMDAN */

static int iceemdan_impl(const double* y, int N, int I, int M, int stop_mode, int S_or_cap,
                         int noise_cap, double eps0, const double* noise, int B, int H,
                         int nthreads, double* imfs, int* trace, int trace_passes, int* stops,
                         double* noise_modes_out, double* r_hist);

int ocm_iceemdan(const double* y, int N, int I, int M, int stop_mode, int S_or_cap,
                 int noise_cap, double eps0, const double* noise, int B, int H, int nthreads,
                 double* imfs, int* trace, int trace_passes, int* stops, double* noise_modes_out)
{
    return iceemdan_impl(y, N, I, M, stop_mode, S_or_cap, noise_cap, eps0, noise, B, H, nthreads,
                         imfs, trace, trace_passes, stops, noise_modes_out, NULL);
}

/* r_hist (optional): M x N, row 0 = x~, row k = r_k (stage k input) */
static int iceemdan_impl(const double* y, int N, int I, int M, int stop_mode, int S_or_cap,
                         int noise_cap, double eps0, const double* noise, int B, int H,
                         int nthreads, double* imfs, int* trace, int trace_passes, int* stops,
                         double* noise_modes_out, double* r_hist)
{
    if (N < 4 || I < 1 || M < 2 || S_or_cap < 1 || noise_cap < 1 || B < 3 ||
        H < 1 || 2L * H >= (long)NLANES * B)
        return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    long IN = (long)I * N;
    double* modes = (double*)malloc(sizeof(double) * (size_t)(M - 1) * IN);
    double* wres = (double*)malloc(sizeof(double) * IN);
    double* xs = (double*)malloc(sizeof(double) * IN);   /* noised signals x_i */
    double* hs = (double*)malloc(sizeof(double) * IN);   /* exported modes h_i */
    double* r = (double*)malloc(sizeof(double) * N);
    double* xt = (double*)malloc(sizeof(double) * N);
    double* inv_i = (double*)malloc(sizeof(double) * I);
    if (!modes || !wres || !xs || !hs || !r || !xt || !inv_i) return -2;
    memcpy(wres, noise, sizeof(double) * IN);

    /* noise modes: E_{k+1}(w^i), k = 0..M-2 (cudaICEEMDAN.cu:845-994).  Each
     * realization is independent, so the loop over i is outermost. */
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < I; ++i) {
        double* work = (double*)malloc(sizeof(double) * 3 * N);
        double* w = wres + (long)i * N;
        for (int k = 0; k < M - 1; ++k) {
            double* E = modes + (long)k * IN + (long)i * N;
            int* tr = NULL;
            if (trace) tr = trace + ((((long)0 * (M - 1) + k) * trace_passes) * I) * 2;
            int tmp[2 * 4096];
            int tp = trace_passes < 4096 ? trace_passes : 4096;
            int subs = ocm_sift_mode(w, N, stop_mode, stop_mode == OCM_STOP_FIXED ? S_or_cap : noise_cap,
                                     B, H, E, work, tr ? tmp : NULL, tp);
            if (tr)
                for (int j = 0; j < tp; ++j) {
                    tr[((long)j * I + i) * 2] = tmp[2 * j];
                    tr[((long)j * I + i) * 2 + 1] = tmp[2 * j + 1];
                }
            if (stops) stops[((long)0 * (M - 1) + k) * I + i] = subs;
            for (int p = 0; p < N; ++p) w[p] = w[p] - E[p];
        }
        free(work);
    }

    /* stage 0: x~ = x / std(x); x_i = x~ + eps0 * E_1^i / std(E_1^i)
     * (cudaICEEMDAN.cu:1011-1024, kernels :57-78) */
    double var_x = ocm_var_canon(y, N);
    double inv_x = 1.0 / sqrt(var_x);
    for (int p = 0; p < N; ++p) xt[p] = y[p] * inv_x;
    if (r_hist) memcpy(r_hist, xt, sizeof(double) * N);
    for (int i = 0; i < I; ++i) inv_i[i] = 1.0 / sqrt(ocm_var_canon(modes + (long)i * N, N));

    for (int k = 0; k < M - 1; ++k) {
        double sq = 0.0;
        if (k > 0) sq = sqrt(ocm_var_canon(r, N));
        const double* Ek = modes + (long)k * IN;
#pragma omp parallel for schedule(dynamic, 1)
        for (int i = 0; i < I; ++i) {
            double* work = (double*)malloc(sizeof(double) * 3 * N);
            double* x_i = xs + (long)i * N;
            const double* E = Ek + (long)i * N;
            if (k == 0) {
                for (int p = 0; p < N; ++p) x_i[p] = xt[p] + eps0 * E[p] * inv_i[i];
            } else {
                for (int p = 0; p < N; ++p) x_i[p] = r[p] + eps0 * E[p] * sq;
            }
            int* tr = NULL;
            if (trace) tr = trace + ((((long)1 * (M - 1) + k) * trace_passes) * I) * 2;
            int tmp[2 * 4096];
            int tp = trace_passes < 4096 ? trace_passes : 4096;
            int subs = ocm_sift_mode(x_i, N, stop_mode, S_or_cap, B, H, hs + (long)i * N, work,
                                     tr ? tmp : NULL, tp);
            if (tr)
                for (int j = 0; j < tp; ++j) {
                    tr[((long)j * I + i) * 2] = tmp[2 * j];
                    tr[((long)j * I + i) * 2 + 1] = tmp[2 * j + 1];
                }
            if (stops) stops[((long)1 * (M - 1) + k) * I + i] = subs;
            free(work);
        }
        /* ensemble mean of the local means (kernels :41-55, :582-607),
         * summed sequentially over realizations for every sample */
        double* imf = imfs + (long)k * N;
        if (k == 0) {
            double invI = 1.0 / (double)I;
            for (int p = 0; p < N; ++p) {
                double tem = 0.0;
                for (int i = 0; i < I; ++i)
                    tem = tem + invI * (xs[(long)i * N + p] - hs[(long)i * N + p]);
                r[p] = tem;
                imf[p] = xt[p] - tem;
            }
        } else {
            for (int p = 0; p < N; ++p) {
                double tem = 0.0;
                for (int i = 0; i < I; ++i) tem = tem + (xs[(long)i * N + p] - hs[(long)i * N + p]);
                imf[p] = r[p] - tem / (double)I;
                r[p] = tem / (double)I;
            }
        }
        if (r_hist) memcpy(r_hist + (long)(k + 1) * N, r, sizeof(double) * N);
    }
    memcpy(imfs + (long)(M - 1) * N, r, sizeof(double) * N);
    if (noise_modes_out) memcpy(noise_modes_out, modes, sizeof(double) * (size_t)(M - 1) * IN);
    free(modes); free(wres); free(xs); free(hs); free(r); free(xt); free(inv_i);
    return 0;
}

/* ------------------------------------------------------------------ Philox */

static inline void mulhilo64(uint64_t a, uint64_t b, uint64_t* lo, uint64_t* hi)
{
    unsigned __int128 p = (unsigned __int128)a * (unsigned __int128)b;
    *lo = (uint64_t)p;
    *hi = (uint64_t)(p >> 64);
}

void ocm_philox4x64(const uint64_t ctr_in[4], const uint64_t key_in[2], uint64_t out[4])
{
    const uint64_t M0 = 0xD2E7470EE14C6C93ULL, M1 = 0xCA5A826395121157ULL;
    const uint64_t W0 = 0x9E3779B97F4A7C15ULL, W1 = 0xBB67AE8584CAA73BULL;
    uint64_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint64_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t lo0, hi0, lo1, hi1;
        mulhilo64(M0, c0, &lo0, &hi0);
        mulhilo64(M1, c2, &lo1, &hi1);
        uint64_t n0 = hi1 ^ c1 ^ k0;
        uint64_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += W0; k1 += W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Gaussian stream: value (i, p) comes from block q = p / 4 with counter
 * (q, i, 0, 0) and key (seed, 0); uniforms u = ((bits >> 11) + 0.5) * 2^-53;
 * Box-Muller on the pairs (u0, u1) and (u2, u3). */
void ocm_noise(uint64_t seed, long I, long N, double* out)
{
    const double TWO_PI = 6.283185307179586476925286766559;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < I; ++i) {
        for (long q = 0; 4 * q < N; ++q) {
            uint64_t ctr[4] = {(uint64_t)q, (uint64_t)i, 0, 0};
            uint64_t key[2] = {seed, 0};
            uint64_t bits[4];
            ocm_philox4x64(ctr, key, bits);
            double u[4];
            for (int j = 0; j < 4; ++j) u[j] = ((double)(bits[j] >> 11) + 0.5) * 0x1.0p-53;
            double z[4];
            double r0 = sqrt(-2.0 * log(u[0]));
            double r1 = sqrt(-2.0 * log(u[2]));
            z[0] = r0 * cos(TWO_PI * u[1]);
            z[1] = r0 * sin(TWO_PI * u[1]);
            z[2] = r1 * cos(TWO_PI * u[3]);
            z[3] = r1 * sin(TWO_PI * u[3]);
            for (int j = 0; j < 4 && 4 * q + j < N; ++j) out[i * N + 4 * q + j] = z[j];
        }
    }
}

/* ================================================================ adjoint
 * With the extrema sets of every pass frozen, one sift pass is linear:
 * h_{j+1} = h_j - (U(h_j) + L(h_j)) / 2, and an envelope is the chain
 *   knot gather -> endpoint fix -> rhs f(a) -> c = A^{-1} f
 *   -> interval coefficients (a, b, c, d) -> evaluation at every sample.
 * The adjoint runs the chain backwards (scatter-add for the gather, A^T
 * solve, transposed evaluation as per-interval moments of the gradient).
 * The A^T solve here is the global Thomas algorithm (exact). */

/* Thomas solve of A^T x = r for the natural-spline matrix A of the knots. */
static void solve_transposed(int n, const int* pos, const double* r, double* x)
{
    double* cp = (double*)malloc(sizeof(double) * n * 2);
    double* dp = cp + n;
    for (int g = 0; g < n; ++g) {
        /* row g of A^T: sub = A[g-1][g], diag = A[g][g], sup = A[g+1][g] */
        double sub = 0.0, sup = 0.0, diag = 1.0;
        if (g >= 1 && g - 1 >= 1 && g - 1 <= n - 2) sub = (double)(pos[g] - pos[g - 1]);
        if (g >= 1 && g <= n - 2) diag = (double)(2 * ((pos[g] - pos[g - 1]) + (pos[g + 1] - pos[g])));
        if (g + 1 <= n - 2 && g + 1 >= 1) sup = (double)(pos[g + 1] - pos[g]);
        if (g == 0) {
            cp[g] = sup / diag;
            dp[g] = r[g] / diag;
        } else {
            double den = diag - sub * cp[g - 1];
            cp[g] = sup / den;
            dp[g] = (r[g] - sub * dp[g - 1]) / den;
        }
    }
    x[n - 1] = dp[n - 1];
    for (int g = n - 2; g >= 0; --g) x[g] = dp[g] - cp[g] * x[g + 1];
    free(cp);
}

void ocm_envelope_adjoint(int N, int n, const int* pos, int fixed_lo, int fixed_hi,
                          const double* g, double* out)
{
    double* mo = (double*)calloc((size_t)4 * n + 3 * (size_t)n, sizeof(double));
    double* M0 = mo; double* M1 = mo + n; double* M2 = mo + 2 * n; double* M3 = mo + 3 * n;
    double* gc = mo + 4 * n; double* gf = gc + n; double* ga = gf + n;
    /* per-interval moments of g (transposed evaluation) */
    int k = 0;
    for (int p = 0; p < N; ++p) {
        while (k + 1 < n - 1 && p > pos[k + 1]) ++k;
        double t = (double)(p - pos[k]);
        M0[k] += g[p];
        M1[k] += g[p] * t;
        M2[k] += g[p] * t * t;
        M3[k] += g[p] * t * t * t;
    }
    /* coefficient formulas (cudaICEEMDAN.cu:394-398) transposed */
    for (int q = 0; q <= n - 2; ++q) {
        double hq = (double)(pos[q + 1] - pos[q]);
        ga[q] += M0[q];
        gc[q] += M2[q];
        /* b_q = (a_{q+1}-a_q)/h - h (2 c_q + c_{q+1}) / 3 */
        ga[q + 1] += M1[q] / hq;
        ga[q] -= M1[q] / hq;
        gc[q] -= 2.0 * hq / 3.0 * M1[q];
        gc[q + 1] -= hq / 3.0 * M1[q];
        /* d_q = (c_{q+1} - c_q) / (3 h) */
        gc[q + 1] += M3[q] / (3.0 * hq);
        gc[q] -= M3[q] / (3.0 * hq);
    }
    /* c = A^{-1} f  ->  gf = A^{-T} gc */
    solve_transposed(n, pos, gc, gf);
    /* f_g = 3 ((a_{g+1}-a_g)/h_g - (a_g-a_{g-1})/h_{g-1}), interior g */
    for (int q = 1; q <= n - 2; ++q) {
        double hl = (double)(pos[q] - pos[q - 1]);
        double hr = (double)(pos[q + 1] - pos[q]);
        ga[q + 1] += 3.0 * gf[q] / hr;
        ga[q] -= 3.0 * gf[q] / hr + 3.0 * gf[q] / hl;
        ga[q - 1] += 3.0 * gf[q] / hl;
    }
    /* endpoint fix (cudaICEEMDAN.cu:209-291) with its forward branch */
    if (n >= 6 && fixed_lo) {
        double rr = (double)(pos[1] - pos[0]) / (double)(pos[2] - pos[1]);
        ga[1] += ga[0] * (1.0 + rr);
        ga[2] -= ga[0] * rr;
        ga[0] = 0.0;
    }
    if (n >= 6 && fixed_hi) {
        double rr = (double)(pos[n - 1] - pos[n - 2]) / (double)(pos[n - 2] - pos[n - 3]);
        ga[n - 2] += ga[n - 1] * (1.0 + rr);
        ga[n - 3] -= ga[n - 1] * rr;
        ga[n - 1] = 0.0;
    }
    for (int q = 0; q < n; ++q) out[pos[q]] += ga[q];
    free(mo);
}

/* forward of one mode extraction recording the frozen structure of each pass */
typedef struct {
    int nmax, nmin, fix[4];
    int* pmax;
    int* pmin;
} pass_rec;

static void sift_record(const double* x, int N, int S, int B, int H, double* mode, pass_rec* rec)
{
    double* h = (double*)malloc(sizeof(double) * N * 2);
    double* hn = h + N;
    double* val = (double*)malloc(sizeof(double) * N * 4);
    double* c = val + N; double* U = c + N; double* L = U + N;
    int* pos = (int*)malloc(sizeof(int) * N);
    memcpy(h, x, sizeof(double) * N);
    for (int j = 0; j < S; ++j) {
        for (int ty = 0; ty < 2; ++ty) {
            int n = ocm_extrema(h, N, ty == 0, pos, val);
            double v0 = val[0], v1 = val[n - 1];
            ocm_boundary_fix(n, pos, val, ty == 0);
            rec[j].fix[2 * ty] = (n >= 6) && (val[0] != v0);
            rec[j].fix[2 * ty + 1] = (n >= 6) && (val[n - 1] != v1);
            int* keep = (int*)malloc(sizeof(int) * n);
            memcpy(keep, pos, sizeof(int) * n);
            if (ty == 0) { rec[j].nmax = n; rec[j].pmax = keep; }
            else { rec[j].nmin = n; rec[j].pmin = keep; }
            ocm_spline_solve(n, pos, val, B, H, c);
            ocm_envelope(N, n, pos, val, c, ty == 0 ? U : L);
        }
        int use_mean = rec[j].nmax > 3 && rec[j].nmin > 3;
        for (int p = 0; p < N; ++p) hn[p] = h[p] - (use_mean ? (U[p] + L[p]) / 2.0 : 0.0);
        memcpy(h, hn, sizeof(double) * N);
    }
    memcpy(mode, h, sizeof(double) * N);
    free(h); free(val); free(pos);
}

/* mu <- Q^T mu for the recorded passes (Q = P_{S-1} ... P_0) */
static void sift_adjoint(int N, int S, const pass_rec* rec, double* mu, double* tmp)
{
    for (int j = S - 1; j >= 0; --j) {
        if (!(rec[j].nmax > 3 && rec[j].nmin > 3)) continue;
        for (int p = 0; p < N; ++p) tmp[p] = 0.0;
        ocm_envelope_adjoint(N, rec[j].nmax, rec[j].pmax, rec[j].fix[0], rec[j].fix[1], mu, tmp);
        ocm_envelope_adjoint(N, rec[j].nmin, rec[j].pmin, rec[j].fix[2], rec[j].fix[3], mu, tmp);
        for (int p = 0; p < N; ++p) mu[p] = mu[p] - 0.5 * tmp[p];
    }
}

int ocm_iceemdan_grad(const double* y, int N, int I, int M, int S, double eps0,
                      const double* noise, int B, int H, int nthreads, const double* grad_imfs,
                      double* grad_y, double* imfs)
{
    if (N < 4 || I < 1 || M < 2 || S < 1) return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    long IN = (long)I * N;
    double* nm = (double*)malloc(sizeof(double) * (size_t)(M - 1) * IN);
    double* rs = (double*)malloc(sizeof(double) * (size_t)M * N);    /* r_0 = x~, r_1 .. r_{M-1} */
    double* gr = (double*)calloc((size_t)M * N, sizeof(double));     /* dL/dr_k, gr[0] = dL/dx~ */
    int rc = iceemdan_impl(y, N, I, M, OCM_STOP_FIXED, S, S, eps0, noise, B, H, nthreads, imfs,
                           NULL, 0, NULL, nm, rs);
    if (rc) return rc;
    double var_x = ocm_var_canon(y, N);
    double s_x = 1.0 / sqrt(var_x);
    /* IMF_0 = x~ - r_1, IMF_k = r_k - r_{k+1}, IMF_{M-1} = r_{M-1} */
    for (int p = 0; p < N; ++p) {
        gr[p] += grad_imfs[p];
        gr[N + p] -= grad_imfs[p];
    }
    for (int k = 1; k <= M - 2; ++k)
        for (int p = 0; p < N; ++p) {
            gr[(long)k * N + p] += grad_imfs[(long)k * N + p];
            gr[(long)(k + 1) * N + p] -= grad_imfs[(long)k * N + p];
        }
    for (int p = 0; p < N; ++p) gr[(long)(M - 1) * N + p] += grad_imfs[(long)(M - 1) * N + p];

    double* inv_i = (double*)malloc(sizeof(double) * I);
    for (int i = 0; i < I; ++i) inv_i[i] = 1.0 / sqrt(ocm_var_canon(nm + (long)i * N, N));
    double* gsum = (double*)malloc(sizeof(double) * (size_t)I * N);
    double* dots = (double*)malloc(sizeof(double) * I);
    for (int k = M - 2; k >= 0; --k) {
        const double* rk = rs + (long)k * N;
        const double* lam = gr + (long)(k + 1) * N;
        double sq = k > 0 ? sqrt(ocm_var_canon(rk, N)) : 0.0;
#pragma omp parallel for schedule(dynamic, 1)
        for (int i = 0; i < I; ++i) {
            const double* E = nm + (long)k * IN + (long)i * N;
            double* xi = (double*)malloc(sizeof(double) * N * 4);
            double* mode = xi + N; double* mu = mode + N; double* tmp = mu + N;
            for (int p = 0; p < N; ++p)
                xi[p] = k == 0 ? rk[p] + eps0 * E[p] * inv_i[i] : rk[p] + eps0 * E[p] * sq;
            pass_rec* rec = (pass_rec*)calloc(S, sizeof(pass_rec));
            sift_record(xi, N, S, B, H, mode, rec);
            memcpy(mu, lam, sizeof(double) * N);
            sift_adjoint(N, S, rec, mu, tmp);
            /* d r_{k+1} / d x_i = (I - Q_i) / I */
            double* gx = gsum + (long)i * N;
            double dot = 0.0;
            for (int p = 0; p < N; ++p) {
                gx[p] = (lam[p] - mu[p]) / (double)I;
                dot += gx[p] * E[p];
            }
            dots[i] = dot;
            for (int j = 0; j < S; ++j) { free(rec[j].pmax); free(rec[j].pmin); }
            free(rec);
            free(xi);
        }
        double* g = gr + (long)k * N;
        double dsum = 0.0;
        for (int i = 0; i < I; ++i) dsum += dots[i];
        for (int p = 0; p < N; ++p) {
            double acc = 0.0;
            for (int i = 0; i < I; ++i) acc += gsum[(long)i * N + p];
            g[p] += acc;
        }
        if (k > 0) {
            /* x_i = r_k + eps0 sigma(r_k) E_i: chain through sigma(r_k) */
            double mu_r = ocm_sum_canon(rk, N) * (1.0 / (double)N);
            double coef = eps0 * dsum / ((double)(N - 1) * sq);
            for (int p = 0; p < N; ++p) g[p] += coef * (rk[p] - mu_r);
        }
    }
    /* x~ = y / std(y) */
    double mu_y = ocm_sum_canon(y, N) * (1.0 / (double)N);
    double dot = 0.0;
    for (int p = 0; p < N; ++p) dot += gr[p] * y[p];
    for (int p = 0; p < N; ++p)
        grad_y[p] = s_x * gr[p] - s_x * s_x * s_x * dot * (y[p] - mu_y) / (double)(N - 1);
    free(nm); free(rs); free(gr); free(inv_i); free(gsum); free(dots);
    return 0;
}
