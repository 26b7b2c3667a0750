/*
 * CPU oracle for the FMM-FFT hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker: the pytest suite, the smoke test and the
 * CPU-baseline leg of bench.py load it to compute expected results.  The
 * product (the HIP library under fmm-fft_amd/) never links or calls it.
 *
 * It restates, in plain C99 and FP64, the algorithm of the reference
 * src/fmmfft.c (97amarnathk/fmm-fft) with the P MPI ranks folded into one
 * address space: P blocks, one global periodic tree, the reference's
 * operators, interaction lists, near/far split and half-shift rule, the
 * fo-combine and a length-M FFT per block.  It also holds an FP64 FFT with
 * long-double twiddles (the "FFTW FP64 oracle" role; FFTW is not installed),
 * a direct long-double DFT, and a brute-force cot-kernel sum used to pin the
 * fast multipole part.  Every function cites the reference lines it follows.
 */
#ifndef FMMFFT_ORACLE_H
#define FMMFFT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Indices of the eight per-level translation matrices (the reference's
 * shftfn, shftfp, flip2n, flip2p, flip3n, flip3p, shftln, shftlp,
 * src/fmmfft.c:281-290). */
enum {
    ORC_SHFTFN = 0, ORC_SHFTFP, ORC_FLIP2N, ORC_FLIP2P,
    ORC_FLIP3N, ORC_FLIP3P, ORC_SHFTLN, ORC_SHFTLP, ORC_NOPS
};

typedef struct orc_plan {
    int64_t N, P, M, terms, b, s, n;
    double *cheb;    /* [terms] Chebyshev nodes                          */
    double *initch;  /* [b][terms] source -> multipole (S2M)             */
    double *ops;     /* [n-2][8][terms][terms] level operators, slot L    */
    double *topflp;  /* [terms][terms] level-2 flip                       */
    double *evalm;   /* [P-1][b][terms] local -> target rows, 1/M scaled  */
    double *evalmh;  /* [P/2][terms] row for target b of the previous box */
    double *kv;      /* [P-1][4b] near kernel K_sc(d), d = -2b .. 2b-1    */
    double *fo;      /* [P][P-1] complex (re,im) combine factors          */
} orc_plan;

/* Returns 0 on success, -1 on invalid (N, P, terms, b). */
int  orc_plan_init(orc_plan *pl, int64_t N, int64_t P, int64_t terms, int64_t b);
void orc_plan_free(orc_plan *pl);

/* Tree storage helpers: levels 2..n, level l has 2^l nodes, each node holds
 * [P-1][terms] complex values (interleaved re,im doubles). */
int64_t orc_tree_len(const orc_plan *pl);           /* complex values */
int64_t orc_level_off(const orc_plan *pl, int l);   /* complex offset */

/* FMM stage by stage.  x, psi_sec, y, X are complex (re,im interleaved).
 *  x       [M][P]  natural-order input (x[mu*P+sc])
 *  sumsec  [P]     (i/M) * sum of section sc (entry 0 unused)
 *  phi,psi tree arrays (orc_tree_len complex each)
 *  psi_sec [M][P]  cot-sum of section sc at target mu (column 0 unused)
 *  y       [P][M]  combined blocks before the local FFTs
 *  X       [N]     final transform                                       */
void orc_section_sums(const orc_plan *pl, const double *x, double *sumsec);
void orc_upward(const orc_plan *pl, const double *x, double *phi);
void orc_downward(const orc_plan *pl, const double *phi, double *psi);
void orc_evaluate(const orc_plan *pl, const double *x, const double *psi, double *psi_sec);
void orc_combine(const orc_plan *pl, const double *x, const double *psi_sec,
                 const double *sumsec, double *y);
void orc_fmmfft(const orc_plan *pl, const double *x, double *X, int dir);

/* FP64 FFT, in place, length n (power of two); sign -1 forward, +1 inverse
 * (unnormalised).  Twiddles are computed in long double. */
void orc_fft64(int64_t n, double *data, int sign);
/* Direct O(n^2) DFT accumulated in long double; sign as above. */
void orc_dft_direct(int64_t n, const double *in, double *out, int sign);
/* Brute-force periodic cot sum: out[mu] = sum_nu K(nu-mu) x_sc[nu] with
 * K(d) = -cot(pi (d + sc/P) / M) / M over all nu (the exact psi_sec). */
void orc_cot_direct(int64_t M, int64_t P, int64_t sc, const double *x,
                    int64_t stride, double *out);

/* Error metrics against a reference y (complex, length n). */
double orc_rel_l2(int64_t n, const double *x, const double *y);
double orc_max_abs(int64_t n, const double *x, const double *y);
/* max_i |x_i - y_i| / |y_i| * 100, the reference's maxError. */
double orc_max_rel_pct(int64_t n, const double *x, const double *y);

/* Deterministic inputs shared with the GPU tests (splitmix64 stream). */
void orc_fill_uniform(int64_t n, uint64_t seed, double lo, double hi, double *out);
void orc_fill_normal(int64_t n, uint64_t seed, double sigma, double *out);
void orc_fill_ramp(int64_t n, double *out);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
