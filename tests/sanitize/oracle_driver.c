/* Driver of the FP64 oracle for sanitizer builds (test infrastructure):
 * every oracle entry point on small inputs, including the edge cases the
 * tests use (few knots, NaN and plateaus, RILLING caps of one pass,
 * not-a-knot ends, the adjoint).  Built with -fsanitize=address,undefined
 * together with oracle/iceemdan_oracle.c; exits 0 when every call returned
 * and no sanitizer reported. */
#define _DEFAULT_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "iceemdan_oracle.h"

static double* signal_of(int N, int kind)
{
    double* x = malloc(sizeof(double) * N);
    for (int p = 0; p < N; ++p) {
        const double t = p / 1000.0;
        x[p] = sin(2 * M_PI * 5 * t) + 0.5 * sin(2 * M_PI * 40 * t) + 0.01 * ((p * 7919) % 13);
        if (kind == 1 && p % 97 == 5) x[p] = x[p - 1];          /* plateaus */
        if (kind == 2 && p == N / 3) x[p] = NAN;                /* a NaN */
    }
    return x;
}

int main(void)
{
    const int N = 1500, I = 4, M = 4;
    double* noise = malloc(sizeof(double) * I * N);
    ocm_noise(7, I, N, noise);
    double* imfs = malloc(sizeof(double) * M * N);
    double* modes = malloc(sizeof(double) * (M - 1) * I * N);
    int* trace = malloc(sizeof(int) * 2 * (M - 1) * 3 * I * 2);
    int* stops = malloc(sizeof(int) * 2 * (M - 1) * I);
    for (int kind = 0; kind < 3; ++kind) {
        double* x = signal_of(N, kind);
        int* pos = malloc(sizeof(int) * N);
        double* val = malloc(sizeof(double) * N);
        double* c = malloc(sizeof(double) * N);
        double* env = malloc(sizeof(double) * N);
        double* hn = malloc(sizeof(double) * N);
        for (int is_max = 0; is_max < 2; ++is_max) {
            const int n = ocm_extrema(x, N, is_max, pos, val);
            ocm_boundary_fix(n, pos, val, is_max);
            for (int bc = 0; bc < 2; ++bc) {
                ocm_set_boundary(bc);
                ocm_spline_solve(n, pos, val, 8, 32, c);
                ocm_spline_solve_thomas(n, pos, val, c);
                ocm_envelope(N, n, pos, val, c, env);
            }
            ocm_set_boundary(OCM_BC_NATURAL);
            double* out = calloc(N, sizeof(double));
            ocm_envelope_adjoint(N, n, pos, 1, 1, env, out);
            free(out);
        }
        int nmax, nmin, mark;
        int64_t cnt;
        ocm_sift_pass(x, N, 8, 32, hn, &nmax, &nmin, &cnt, &mark);
        (void)ocm_rilling_continue(cnt, mark, nmax, nmin, N);
        for (int stop = 0; stop < 2; ++stop)
            for (int cap = 1; cap <= 12; cap += 11)
                if (ocm_iceemdan(x, N, I, M, stop, cap, cap, 0.2, noise, 8, 32, 2, imfs, trace, 3,
                                 stops, modes) != 0)
                    return 2;
        double* G = malloc(sizeof(double) * M * N);
        double* gy = malloc(sizeof(double) * N);
        for (int q = 0; q < M * N; ++q) G[q] = cos(q / 17.0);
        if (kind != 2 &&
            ocm_iceemdan_grad(x, N, I, M, 6, 0.2, noise, 8, 32, 2, G, gy, imfs) != 0)
            return 3;
        free(G); free(gy);
        free(x); free(pos); free(val); free(c); free(env); free(hn);
    }
    /* tiny series: 4 samples, fewer than four knots */
    double tiny[4] = {0.0, 1.0, -1.0, 0.5};
    double tn[4 * 1];
    ocm_noise(3, 1, 4, tn);
    double ti[2 * 4];
    if (ocm_iceemdan(tiny, 4, 1, 2, 0, 3, 3, 0.2, tn, 8, 32, 1, ti, NULL, 0, NULL, NULL) != 0)
        return 4;
    free(noise); free(imfs); free(modes); free(trace); free(stops);
    printf("oracle driver ok\n");
    return 0;
}
