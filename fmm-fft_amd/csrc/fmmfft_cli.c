/* Command-line driver with the reference's interface and output line.
 *
 *   fmmfft_amd N B T [P] [--c64] [--runs R]
 *
 * N points, B points per leaf box, T Chebyshev terms, P blocks (the
 * reference takes P from the MPI world size, `mpiexec -n P fmmfft.out N B T`,
 * src/fmmfft.c:1192-1212; default 4 here).  Input is the reference's ramp
 * x_n = n + 1 (src/fmmfft.c:1228-1246).  The plan is built once and the
 * transform run R times (default 6, src/fmmfft.c:1249-1275); every run
 * prints the reference's CSV line
 *
 *   N, P, B, T, time_s, err_pct
 *
 * (src/fmmfft.c:1266) where err_pct is a valid sixth column: the error
 * metric of src/fmmfft_error.c:1097-1108,1267, max over the first N/P
 * outputs of |X_k - Y_k| / |Y_k| * 100, with Y the exact DFT of the ramp,
 * Y_0 = N(N+1)/2 and Y_k = N / (e^{-2 pi i k/N} - 1) for k > 0 (from
 * sum_n n z^n = N / (z - 1) for z^N = 1, z != 1), evaluated in long double
 * (the reference compares against FFTW's FP64 output instead).  time_s is
 * the execute time from HIP events on the execute stream (device-resident
 * input; the ramp is copied once).
 */
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/fmmfft.h"

static void usage(const char *argv0) {
    fprintf(stderr,
            "usage: %s N B T [P] [--c64] [--runs R]\n"
            "  N points (power of two), B points per leaf box, T Chebyshev terms,\n"
            "  P blocks (power of two, 2..64, default 4); --c64 runs the complex64\n"
            "  precision form (default complex128); R timed runs (default 6).\n"
            "Prints 'N, P, B, T, time_s, err_pct' per run.\n",
            argv0);
}

static int parse_long(const char *s, long *out) {
    char *end = NULL;
    long v = strtol(s, &end, 10);
    if (end == s || *end != '\0') return 0;
    *out = v;
    return 1;
}

/* max_k<m |X_k - Y_k| / |Y_k| * 100 against the exact DFT of x_n = n + 1 */
static double ramp_max_err_pct(const double *X, long n, long m, int c64, const float *Xf) {
    const long double two_pi = 6.283185307179586476925286766559005768L;
    double worst = 0.0;
    for (long k = 0; k < m; ++k) {
        long double yr, yi;
        if (k == 0) {
            yr = (long double)n * ((long double)n + 1.0L) / 2.0L;
            yi = 0.0L;
        } else {
            /* e^{-ia} - 1 = -2 sin^2(a/2) - i sin(a): no cancellation at small a */
            const long double a = two_pi * (long double)k / (long double)n;
            const long double h = sinl(a / 2.0L);
            const long double dr = -2.0L * h * h, di = -sinl(a);
            const long double den = dr * dr + di * di;
            yr = (long double)n * dr / den;
            yi = -(long double)n * di / den;
        }
        const long double xr = c64 ? (long double)Xf[2 * k] : (long double)X[2 * k];
        const long double xi = c64 ? (long double)Xf[2 * k + 1] : (long double)X[2 * k + 1];
        const long double er = xr - yr, ei = xi - yi;
        const double e = (double)(sqrtl(er * er + ei * ei) / sqrtl(yr * yr + yi * yi) * 100.0L);
        if (e > worst || k == 0) worst = e;
    }
    return worst;
}

#define HIP_OK(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            fprintf(stderr, "%s failed: %s\n", #call, hipGetErrorString(e_));          \
            return 1;                                                                  \
        }                                                                              \
    } while (0)

int main(int argc, char **argv) {
    long pos[4] = {0, 0, 0, 4};
    int npos = 0, c64 = 0;
    long runs = 6;
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "--c64")) {
            c64 = 1;
        } else if (!strcmp(argv[i], "--runs") && i + 1 < argc) {
            if (!parse_long(argv[++i], &runs) || runs < 1) {
                usage(argv[0]);
                return 2;
            }
        } else if (!strcmp(argv[i], "-h") || !strcmp(argv[i], "--help")) {
            usage(argv[0]);
            return 0;
        } else if (npos < 4 && parse_long(argv[i], &pos[npos])) {
            ++npos;
        } else {
            usage(argv[0]);
            return 2;
        }
    }
    if (npos < 3) {
        /* the reference prints this and exits (src/fmmfft.c:1200-1204) */
        printf("ERROR : Provide correct args\n");
        usage(argv[0]);
        return 2;
    }
    const long N = pos[0], B = pos[1], T = pos[2], P = pos[3];
    const int dtype = c64 ? FMMFFT_C64 : FMMFFT_C128;

    /* parameter check before any device work */
    size_t tab_bytes = 0, work_bytes = 0;
    int rc = fmmfft_plan_sizes(N, (int)P, (int)T, (int)B, dtype, &tab_bytes, &work_bytes);
    if (rc != FMMFFT_OK) {
        fprintf(stderr, "fmmfft_amd: N=%ld P=%ld B=%ld T=%ld: %s\n", N, P, B, T, fmmfft_strerror(rc));
        return 2;
    }

    const size_t esz = c64 ? 8 : 16;
    void *host = malloc((size_t)N * esz);
    if (!host) {
        fprintf(stderr, "fmmfft_amd: out of host memory\n");
        return 1;
    }
    for (long i = 0; i < N; ++i) {
        if (c64) {
            ((float *)host)[2 * i] = (float)(i + 1);
            ((float *)host)[2 * i + 1] = 0.0f;
        } else {
            ((double *)host)[2 * i] = (double)(i + 1);
            ((double *)host)[2 * i + 1] = 0.0;
        }
    }

    fmmfft_plan plan = NULL;
    rc = fmmfft_plan_create(&plan, N, (int)P, (int)T, (int)B, dtype, 1);
    if (rc != FMMFFT_OK) {
        fprintf(stderr, "fmmfft_amd: plan: %s\n", fmmfft_strerror(rc));
        free(host);
        return 1;
    }
    void *d_in = NULL, *d_out = NULL;
    hipStream_t st;
    hipEvent_t t0, t1;
    HIP_OK(hipMalloc(&d_in, (size_t)N * esz));
    HIP_OK(hipMalloc(&d_out, (size_t)N * esz));
    HIP_OK(hipStreamCreate(&st));
    HIP_OK(hipEventCreate(&t0));
    HIP_OK(hipEventCreate(&t1));
    HIP_OK(hipMemcpy(d_in, host, (size_t)N * esz, hipMemcpyHostToDevice));

    const long M = N / P;
    for (long r = 0; r < runs; ++r) {
        HIP_OK(hipEventRecord(t0, st));
        rc = fmmfft_execute(plan, d_in, d_out, FMMFFT_FORWARD, st);
        HIP_OK(hipEventRecord(t1, st));
        if (rc != FMMFFT_OK) {
            fprintf(stderr, "fmmfft_amd: execute: %s\n", fmmfft_strerror(rc));
            return 1;
        }
        HIP_OK(hipEventSynchronize(t1));
        float ms = 0.0f;
        HIP_OK(hipEventElapsedTime(&ms, t0, t1));
        HIP_OK(hipMemcpy(host, d_out, (size_t)M * esz, hipMemcpyDeviceToHost));
        const double err = ramp_max_err_pct((const double *)host, N, M, c64, (const float *)host);
        printf("%ld, %ld, %ld, %ld, %lf, %lf\n", N, P, B, T, ms * 1e-3, err);
        fflush(stdout);
    }

    fmmfft_plan_destroy(plan);
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    (void)hipStreamDestroy(st);
    free(host);
    return 0;
}
