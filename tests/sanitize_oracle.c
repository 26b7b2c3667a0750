/*
 * Host sanitizer driver for the CPU oracle (test infrastructure only; built
 * and run by tests/test_sanitize.py with clang -fsanitize=address,undefined).
 *
 * Runs every exported oracle routine over small shapes that reach each of
 * its index paths -- P = 2 (no half-shift box), P >= 4 (sections past P/2),
 * the shortest tree (s = 8 leaves), leaf 2 and leaf 64, terms 2..16, both
 * directions -- and checks the results against the oracle's own FFT, so an
 * out-of-bounds access, a use after free, a leak or an undefined operation
 * in any of them stops the run with a sanitizer report.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "fmmfft_oracle.h"

static int fail(const char *what, double v) {
    printf("FAIL %s %.3e\n", what, v);
    return 1;
}

int main(void) {
    static const int64_t cases[][4] = {
        /* N, P, terms, b */
        {1 << 5, 2, 2, 2},     {1 << 8, 2, 8, 16},    {1 << 10, 4, 12, 32},
        {1 << 12, 8, 16, 64},  {1 << 11, 16, 12, 16}, {1 << 12, 64, 16, 8},
        {1 << 13, 32, 20, 32},
    };
    int bad = 0;
    for (size_t c = 0; c < sizeof(cases) / sizeof(cases[0]); ++c) {
        const int64_t N = cases[c][0], P = cases[c][1], terms = cases[c][2], b = cases[c][3];
        orc_plan pl;
        if (orc_plan_init(&pl, N, P, terms, b) != 0) {
            printf("FAIL plan N=%lld P=%lld\n", (long long)N, (long long)P);
            return 1;
        }
        double *x = malloc(sizeof(double) * 2 * N), *X = malloc(sizeof(double) * 2 * N);
        double *F = malloc(sizeof(double) * 2 * N);
        orc_fill_normal(N, 7 + c, 1.0, x);
        for (int dir = -1; dir <= 1; dir += 2) {
            orc_fmmfft(&pl, x, X, dir);
            for (int64_t i = 0; i < 2 * N; ++i) F[i] = x[i];
            orc_fft64(N, F, dir < 0 ? +1 : -1);
            double e = orc_rel_l2(N, X, F);
            /* 2 and 8 terms are coarse expansions (rel-L2 1.5e-2 and 1.1e-8
             * on these inputs); 12 and more reach 1e-12 */
            double tol = terms >= 12 ? 1e-10 : terms >= 8 ? 1e-7 : 0.05;
            if (!(e < tol)) bad |= fail("fmmfft vs fft64", e);
        }
        /* stage by stage (the functions the parity tests read one by one) */
        int64_t tl = orc_tree_len(&pl);
        double *phi = malloc(sizeof(double) * 2 * tl), *psi = malloc(sizeof(double) * 2 * tl);
        double *ps = malloc(sizeof(double) * 2 * N), *ss = malloc(sizeof(double) * 2 * P);
        orc_section_sums(&pl, x, ss);
        orc_upward(&pl, x, phi);
        orc_downward(&pl, phi, psi);
        orc_evaluate(&pl, x, psi, ps);
        orc_combine(&pl, x, ps, ss, X);
        (void)orc_level_off(&pl, (int)pl.n);
        /* brute-force cot sum of the last section against psi_sec's column */
        const int64_t M = N / P, sc = P - 1;
        double *cd = malloc(sizeof(double) * 2 * M);
        orc_cot_direct(M, P, sc, x + 2 * sc, P, cd);
        double num = 0, den = 0;
        for (int64_t mu = 0; mu < M; ++mu)
            for (int k = 0; k < 2; ++k) {
                double d = ps[2 * (mu * P + sc) + k] - cd[2 * mu + k];
                num += d * d;
                den += cd[2 * mu + k] * cd[2 * mu + k];
            }
        if (terms >= 12 && !(sqrt(num / den) < 1e-10)) bad |= fail("psi_sec vs cot_direct", sqrt(num / den));
        free(cd); free(phi); free(psi); free(ps); free(ss);
        free(x); free(X); free(F);
        orc_plan_free(&pl);
    }
    /* the direct DFT and the metric helpers on a short vector */
    {
        const int64_t n = 64;
        double *a = malloc(sizeof(double) * 2 * n), *d = malloc(sizeof(double) * 2 * n),
               *f = malloc(sizeof(double) * 2 * n);
        orc_fill_uniform(n, 3, -1.0, 1.0, a);
        orc_fill_ramp(n, f);
        for (int64_t i = 0; i < 2 * n; ++i) f[i] = a[i];
        orc_dft_direct(n, a, d, -1);
        orc_fft64(n, f, -1);
        double e = orc_rel_l2(n, f, d);
        if (!(e < 1e-14)) bad |= fail("fft64 vs dft_direct", e);
        if (!(orc_max_abs(n, f, d) < 1e-12)) bad |= fail("max_abs", orc_max_abs(n, f, d));
        if (!(orc_max_rel_pct(n, f, d) < 1e-9)) bad |= fail("max_rel_pct", orc_max_rel_pct(n, f, d));
        free(a); free(d); free(f);
    }
    if (orc_num_threads() < 1) bad |= fail("num_threads", 0);
    if (!bad) printf("sanitize ok\n");
    return bad;
}
