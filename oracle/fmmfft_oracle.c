/*
 * CPU oracle for the FMM-FFT hot path -- TEST INFRASTRUCTURE ONLY.
 * Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * as the checker; the product never calls it.
 *
 * A single-address-space restatement of the reference src/fmmfft.c.  The
 * reference spreads the signal over P MPI ranks (rank q owns leaf boxes
 * q*t .. q*t+t-1 of every section, src/fmmfft.c:419-439) and exchanges tree
 * nodes and halo boxes between ranks.  Here every rank's data lives in one
 * array and each exchange becomes a periodic index into a global tree, so
 * the arithmetic (operators, interaction lists, near/far split, half-shift
 * rule, fo-combine, local FFT) is the reference's; only the summation order
 * of a few sums differs.
 *
 * Parity pins (tests/test_oracle.py): every operator table against closed
 * forms and interpolation identities, the FMM section sums against the
 * brute-force cot sum and the fractional-shift identity, the FFT against
 * numpy golden vectors and a long-double DFT, and the whole path against the
 * max-error values the reference recorded in
 * plotting/raw_data/cluster_output_error.csv.
 */
#include "fmmfft_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* src/fmmfft.c:17 */
static const double ORC_PI = 3.14159265358979323846264338327950288;

static int64_t ilog2(int64_t x) {
    int64_t r = 0;
    while (x > 1) { x >>= 1; ++r; }
    return r;
}
static int is_pow2(int64_t x) { return x > 0 && (x & (x - 1)) == 0; }
static int64_t pmod(int64_t a, int64_t m) { a %= m; return a < 0 ? a + m : a; }

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---------------------------------------------------------------------------
 * Operator construction.
 *
 * Every translation is a rational interpolation matrix with prescribed poles:
 *   mat[k][j] = w_j / (x_k - pole_j) * prod_{l != j} (x_k - c_l)/(x_k - pole_l)
 *   w_j       = (c_j - pole_j) * prod_{l != j} (c_j - pole_l)/(c_j - c_l)
 * so that mat[k][j] = delta_kj when x_k = c_k.  Multipole-side matrices use
 * poles 3 tan(a) / tan(th/2), th on Chebyshev-Lobatto points of [-2a, 2a]
 * (src/fmmfft.c:81-114); local-side matrices use cot(a) tan(th/2), th on
 * Chebyshev-Lobatto points of [6a, 2 pi - 6a] (src/fmmfft.c:116-146).
 * ------------------------------------------------------------------------- */

static void interp_matrix(int64_t terms, const double *cheb, const double *pole,
                          int64_t nx, const double *x, double *mat) {
    for (int64_t j = 0; j < terms; ++j) {
        double w = cheb[j] - pole[j];
        for (int64_t l = 0; l < terms; ++l)
            if (l != j) w *= (cheb[j] - pole[l]) / (cheb[j] - cheb[l]);
        for (int64_t k = 0; k < nx; ++k) {
            double v = w / (x[k] - pole[j]);
            for (int64_t l = 0; l < terms; ++l)
                if (l != j) v *= (x[k] - cheb[l]) / (x[k] - pole[l]);
            mat[k * terms + j] = v;
        }
    }
}

/* Multipole-side poles, src/fmmfft.c:82-87. */
static void poles_multipole(double a, int64_t terms, double *pole) {
    double ac = 3.0 * tan(a);
    for (int64_t j = 1; j <= terms; ++j) {
        double th = -1.0 * cos(((double)2 * j - 2) / (2 * (terms - 1)) * ORC_PI) * (2 * a);
        pole[j - 1] = ac / tan(0.5 * th);
    }
}

/* Local-side poles, src/fmmfft.c:117-123. */
static void poles_local(double a, int64_t terms, double *pole) {
    double ac = 1.0 / tan(a);
    for (int64_t j = 1; j <= terms; ++j) {
        double th = -cos(((double)(2 * j - 2)) / (2 * (terms - 1)) * ORC_PI) * (ORC_PI - 6 * a) + ORC_PI;
        pole[j - 1] = ac * tan(0.5 * th);
    }
}

/* M2M shift, src/fmmfft.c:148-159: parent nodes seen from a child of
 * half-width a/2 whose centre is displaced by dir. */
static void op_shift_up(int64_t lev, int dir, int64_t terms, const double *cheb, double *mat) {
    double a = ORC_PI / pow(2, (double)lev);
    double ta3 = 3 * tan(a), td2 = tan(0.5 * a), dd = (double)dir;
    double x[64], pole[64];
    for (int64_t j = 0; j < terms; ++j)
        x[j] = 3 * td2 * (cheb[j] + dd * ta3 * td2) / (ta3 - dd * td2 * cheb[j]);
    poles_multipole(0.5 * a, terms, pole);
    interp_matrix(terms, cheb, pole, terms, x, mat);
}

/* M2L flip, src/fmmfft.c:227-240: a level-lev target box sees a source box
 * shft boxes away. */
static void op_flip(int64_t lev, int shft, int64_t terms, const double *cheb, double *mat) {
    double a = ORC_PI / pow(2, (double)lev);
    double c = tan(0.5 * a), td = tan((double)shft * a);
    double x[64], pole[64];
    for (int64_t j = 0; j < terms; ++j)
        x[j] = 3 * c * (1 - td * c * cheb[j]) / (td + c * cheb[j]);
    poles_multipole(0.5 * a, terms, pole);
    interp_matrix(terms, cheb, pole, terms, x, mat);
}

/* L2L shift, src/fmmfft.c:186-197: child local nodes in the parent's local
 * coordinate. */
static void op_shift_down(int64_t lev, int dir, int64_t terms, const double *cheb, double *mat) {
    double a = ORC_PI / pow(2, (double)lev);
    double c = tan(0.25 * a), td2 = tan(0.5 * a), dd = (double)dir;
    double x[64], pole[64];
    for (int64_t j = 0; j < terms; ++j)
        x[j] = (cheb[j] - dd) / (1 / c + dd * c * cheb[j]) / td2;
    poles_local(0.5 * a, terms, pole);
    interp_matrix(terms, cheb, pole, terms, x, mat);
}

/* L2T row, src/fmmfft.c:161-172: local expansion of a leaf (half-width
 * pi/(2s)) evaluated at box coordinate ffr, scaled by 1/M. */
static void op_eval_row(double ffr, int64_t s, int64_t M, int64_t terms, const double *cheb, double *row) {
    double a = (ORC_PI / (double)s) * 0.5;
    double x = tan(ffr * a) / tan(a);
    double pole[64];
    poles_local(a, terms, pole);
    interp_matrix(terms, cheb, pole, 1, &x, row);
    for (int64_t k = 0; k < terms; ++k) row[k] = row[k] / (double)M;
}

/* S2M, src/fmmfft.c:242-259: cot(theta_k - theta_j) with tan(theta_k) =
 * 3 tan(a) / c_k, i.e. a source's far field at the multipole nodes. */
static void op_s2m(int64_t b, int64_t s, int64_t terms, const double *cheb, double *initch) {
    double a = ORC_PI / (2.0 * (double)s);
    double ta3 = 3 * tan(a);
    for (int64_t j = 1; j <= b; ++j) {
        double tg = tan(((double)(2 * j - b - 1)) / (double)b * a);
        for (int64_t k = 0; k < terms; ++k)
            initch[(j - 1) * terms + k] = (cheb[k] + ta3 * tg) / (ta3 - cheb[k] * tg);
    }
}

int64_t orc_tree_len(const orc_plan *pl) {
    /* levels 2..n: sum 2^l = 2^(n+1) - 4 nodes */
    return (((int64_t)1 << (pl->n + 1)) - 4) * (pl->P - 1) * pl->terms;
}

int64_t orc_level_off(const orc_plan *pl, int l) {
    return (((int64_t)1 << l) - 4) * (pl->P - 1) * pl->terms;
}

int orc_plan_init(orc_plan *pl, int64_t N, int64_t P, int64_t terms, int64_t b) {
    memset(pl, 0, sizeof(*pl));
    if (P < 2 || !is_pow2(P) || N % P || terms < 2 || terms > 64 || b < 2) return -1;
    int64_t M = N / P;
    if (M % b) return -1;
    int64_t s = M / b;
    if (!is_pow2(s) || s < 8) return -1;
    int64_t n = ilog2(s), S = P - 1;
    pl->N = N; pl->P = P; pl->M = M; pl->terms = terms; pl->b = b; pl->s = s; pl->n = n;

    pl->cheb = malloc(sizeof(double) * terms);
    pl->initch = malloc(sizeof(double) * b * terms);
    pl->ops = malloc(sizeof(double) * (n - 2) * ORC_NOPS * terms * terms);
    pl->topflp = malloc(sizeof(double) * terms * terms);
    pl->evalm = malloc(sizeof(double) * S * b * terms);
    pl->evalmh = malloc(sizeof(double) * (P / 2) * terms);
    pl->kv = malloc(sizeof(double) * S * 4 * b);
    pl->fo = malloc(sizeof(double) * 2 * P * S);

    /* Chebyshev nodes, src/fmmfft.c:273-275 */
    for (int64_t j = 1; j <= terms; ++j)
        pl->cheb[j - 1] = -1 * cos(((double)2 * j - 1) / (2 * terms) * ORC_PI);

    op_s2m(b, s, terms, pl->cheb, pl->initch);              /* :277 */
    op_flip(2, 2, terms, pl->cheb, pl->topflp);             /* :279 */
    for (int64_t L = 0; L < n - 2; ++L) {                   /* :281-290 */
        double *o = pl->ops + L * ORC_NOPS * terms * terms;
        int64_t tt = terms * terms;
        op_shift_up(L + 3, -1, terms, pl->cheb, o + ORC_SHFTFN * tt);
        op_shift_up(L + 3, +1, terms, pl->cheb, o + ORC_SHFTFP * tt);
        op_flip(L + 3, -2, terms, pl->cheb, o + ORC_FLIP2N * tt);
        op_flip(L + 3, +2, terms, pl->cheb, o + ORC_FLIP2P * tt);
        op_flip(L + 3, -3, terms, pl->cheb, o + ORC_FLIP3N * tt);
        op_flip(L + 3, +3, terms, pl->cheb, o + ORC_FLIP3P * tt);
        op_shift_down(L + 2, -1, terms, pl->cheb, o + ORC_SHFTLN * tt);
        op_shift_down(L + 2, +1, terms, pl->cheb, o + ORC_SHFTLP * tt);
    }

    double dM = (double)M;
    for (int64_t sc = 1; sc <= S; ++sc) {                   /* :292-321 */
        double gfrac = (double)sc / (double)P;
        for (int64_t j = 1; j <= b; ++j) {
            double ffr = ((double)(2 * j - 1) - ((double)(2 * sc)) / (double)P) / (double)b - 1;
            op_eval_row(ffr, s, M, terms, pl->cheb, pl->evalm + ((sc - 1) * b + (j - 1)) * terms);
        }
        if (sc >= P / 2) {
            double ffr = (((double)2 * b + 1) - ((double)2 * sc) / (double)P) / (double)b - 1;
            op_eval_row(ffr, s, M, terms, pl->cheb, pl->evalmh + (sc - P / 2) * terms);
        }
        /* The reference's cots / cotprv / cotsh tables all sample
         * -1/tan(pi/M (d + sc/P))/M at integer source-target offsets d;
         * kv holds that Toeplitz generator for d = -2b .. 2b-1. */
        for (int64_t d = -2 * b; d < 2 * b; ++d)
            pl->kv[(sc - 1) * 4 * b + (d + 2 * b)] = -1 / tan(ORC_PI / dM * ((double)d + gfrac)) / dM;
    }

    /* combine factors, src/fmmfft.c:325-331 */
    for (int64_t sc = 1; sc <= S; ++sc) {
        double ang = ORC_PI * (double)sc / (double)P;
        for (int64_t j = 1; j <= P; ++j) {
            cplx v = -cexp(I * (ang * (double)(1 - 2 * j))) * sin(ang);
            pl->fo[2 * ((j - 1) * S + (sc - 1))] = creal(v);
            pl->fo[2 * ((j - 1) * S + (sc - 1)) + 1] = cimag(v);
        }
    }
    return 0;
}

void orc_plan_free(orc_plan *pl) {
    free(pl->cheb); free(pl->initch); free(pl->ops); free(pl->topflp);
    free(pl->evalm); free(pl->evalmh); free(pl->kv); free(pl->fo);
    memset(pl, 0, sizeof(*pl));
}

/* ---------------------------------------------------------------------------
 * Execute.
 * ------------------------------------------------------------------------- */

/* Section sums, src/fmmfft.c:441-457 (the per-rank partial sums plus
 * MPI_Reduce/Bcast become one sum over the whole section). */
void orc_section_sums(const orc_plan *pl, const double *x, double *sumsec) {
    const cplx *xc = (const cplx *)x;
    cplx *ss = (cplx *)sumsec;
    ss[0] = 0;
    for (int64_t sc = 1; sc < pl->P; ++sc) {
        cplx acc = 0;
        for (int64_t mu = 0; mu < pl->M; ++mu) acc += xc[mu * pl->P + sc];
        ss[sc] = acc * (I / (double)pl->M);
    }
}

/* out[sc][t] = sum_k mat[t][k] in[sc][k], the reference's mpp
 * (src/fmmfft.c:174-184); accumulate when acc != 0. */
static void apply_op(int64_t S, int64_t terms, const double *mat, const cplx *in, cplx *out, int acc) {
    for (int64_t sc = 0; sc < S; ++sc)
        for (int64_t t = 0; t < terms; ++t) {
            cplx v = 0;
            for (int64_t k = 0; k < terms; ++k) v += in[sc * terms + k] * mat[t * terms + k];
            if (acc) out[sc * terms + t] += v; else out[sc * terms + t] = v;
        }
}

/* S2M (src/fmmfft.c:199-211, 462-465) on every leaf, then M2M up to level 2
 * (src/fmmfft.c:468-515). */
void orc_upward(const orc_plan *pl, const double *x, double *phi_d) {
    const cplx *xc = (const cplx *)x;
    cplx *phi = (cplx *)phi_d;
    int64_t P = pl->P, S = P - 1, T = pl->terms, b = pl->b, n = pl->n;
    cplx *leaf = phi + orc_level_off(pl, (int)n);
    #pragma omp parallel for schedule(static)
    for (int64_t box = 0; box < pl->s; ++box)
        for (int64_t sc = 1; sc <= S; ++sc)
            for (int64_t k = 0; k < T; ++k) {
                cplx acc = 0;
                for (int64_t j = 0; j < b; ++j)
                    acc += xc[(box * b + j) * P + sc] * pl->initch[j * T + k];
                leaf[(box * S + (sc - 1)) * T + k] = acc;
            }
    for (int64_t l = n - 1; l >= 2; --l) {
        const double *o = pl->ops + (l - 2) * ORC_NOPS * T * T;
        cplx *par = phi + orc_level_off(pl, (int)l);
        const cplx *ch = phi + orc_level_off(pl, (int)l + 1);
        int64_t nn = (int64_t)1 << l;
        #pragma omp parallel for schedule(static)
        for (int64_t q = 0; q < nn; ++q) {
            apply_op(S, T, o + ORC_SHFTFN * T * T, ch + (2 * q) * S * T, par + q * S * T, 0);
            apply_op(S, T, o + ORC_SHFTFP * T * T, ch + (2 * q + 1) * S * T, par + q * S * T, 1);
        }
    }
}

/* Top flip (src/fmmfft.c:620-639) and the downward pass with the interaction
 * lists of src/fmmfft.c:644-749: a left child (even index c) takes
 * flip2p(c-2), flip2n(c+2), flip3n(c+3); a right child takes flip3p(c-3),
 * flip2p(c-2), flip2n(c+2); both add the parent's local expansion shifted by
 * shftlp / shftln.  Indices are periodic, as the ring of ranks is. */
void orc_downward(const orc_plan *pl, const double *phi_d, double *psi_d) {
    const cplx *phi = (const cplx *)phi_d;
    cplx *psi = (cplx *)psi_d;
    int64_t S = pl->P - 1, T = pl->terms, n = pl->n, TT = T * T;
    const cplx *p2 = phi + orc_level_off(pl, 2);
    cplx *s2 = psi + orc_level_off(pl, 2);
    for (int64_t q = 0; q < 4; ++q)
        apply_op(S, T, pl->topflp, p2 + ((q + 2) % 4) * S * T, s2 + q * S * T, 0);
    for (int64_t l = 3; l <= n; ++l) {
        const double *o = pl->ops + (l - 3) * ORC_NOPS * TT;
        const cplx *ph = phi + orc_level_off(pl, (int)l);
        const cplx *pp = psi + orc_level_off(pl, (int)l - 1);
        cplx *ps = psi + orc_level_off(pl, (int)l);
        int64_t nn = (int64_t)1 << l;
        #pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < nn; ++c) {
            cplx *out = ps + c * S * T;
            const cplx *par = pp + (c / 2) * S * T;
#define NODE(off) (ph + pmod(c + (off), nn) * S * T)
            if ((c & 1) == 0) {
                apply_op(S, T, o + ORC_SHFTLP * TT, par, out, 0);
                apply_op(S, T, o + ORC_FLIP2P * TT, NODE(-2), out, 1);
                apply_op(S, T, o + ORC_FLIP2N * TT, NODE(+2), out, 1);
                apply_op(S, T, o + ORC_FLIP3N * TT, NODE(+3), out, 1);
            } else {
                apply_op(S, T, o + ORC_SHFTLN * TT, par, out, 0);
                apply_op(S, T, o + ORC_FLIP3P * TT, NODE(-3), out, 1);
                apply_op(S, T, o + ORC_FLIP2P * TT, NODE(-2), out, 1);
                apply_op(S, T, o + ORC_FLIP2N * TT, NODE(+2), out, 1);
            }
#undef NODE
        }
    }
}

/* Far part (L2T) plus near part of one target mu of section sc evaluated as
 * local index jj of home box h.  The near sources are the 3b samples of boxes
 * h-1, h, h+1 (src/fmmfft.c:213-225, 805-954). */
static cplx eval_target(const orc_plan *pl, const cplx *xc, const cplx *psi_leaf,
                        int64_t sc, int64_t h, int64_t mu, const double *row) {
    int64_t P = pl->P, S = P - 1, T = pl->terms, b = pl->b, M = pl->M, s = pl->s;
    int64_t hh = pmod(h, s);
    const cplx *loc = psi_leaf + (hh * S + (sc - 1)) * T;
    cplx far = 0;
    for (int64_t k = 0; k < T; ++k) far += loc[k] * row[k];
    const double *kv = pl->kv + (sc - 1) * 4 * b;
    cplx nr = 0;
    int64_t nu0 = (h - 1) * b;
    for (int64_t q = 0; q < 3 * b; ++q) {
        int64_t nu = nu0 + q;
        nr += xc[pmod(nu, M) * P + sc] * kv[(nu - mu) + 2 * b];
    }
    return far + nr;
}

/* L2T and direct near-field sums with the half-shift rule
 * (src/fmmfft.c:751-954): target mu = i*b + j of section sc sits at mu - sc/P.
 * For j = 0 and sc > P/2 it is evaluated as point b of box i-1 (evalmh,
 * boxes i-2..i); for sc = P/2 the two evaluations are averaged. */
void orc_evaluate(const orc_plan *pl, const double *x, const double *psi_d, double *psi_sec) {
    const cplx *xc = (const cplx *)x;
    const cplx *psi_leaf = (const cplx *)psi_d + orc_level_off(pl, (int)pl->n);
    cplx *out = (cplx *)psi_sec;
    int64_t P = pl->P, T = pl->terms, b = pl->b;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < pl->s; ++i)
        for (int64_t sc = 1; sc < P; ++sc) {
            const double *ev = pl->evalm + ((sc - 1) * b) * T;
            for (int64_t j = 0; j < b; ++j) {
                int64_t mu = i * b + j;
                cplx v;
                if (j == 0 && 2 * sc >= P) {
                    cplx prev = eval_target(pl, xc, psi_leaf, sc, i - 1, mu, pl->evalmh + (sc - P / 2) * T);
                    if (2 * sc == P)
                        v = 0.5 * (eval_target(pl, xc, psi_leaf, sc, i, mu, ev) + prev);
                    else
                        v = prev;
                } else {
                    v = eval_target(pl, xc, psi_leaf, sc, i, mu, ev + j * T);
                }
                out[mu * P + sc] = v;
            }
        }
    for (int64_t mu = 0; mu < pl->M; ++mu) out[mu * P] = 0;
}

/* Combine, src/fmmfft.c:956-968: block j gets x_0 + sum_sc fo[j][sc] *
 * (psi_sc - sumsec_sc); the block becomes the input of the j-th local FFT
 * (the reference's MPI_Alltoall, :970). */
void orc_combine(const orc_plan *pl, const double *x, const double *psi_sec,
                 const double *sumsec, double *y) {
    const cplx *xc = (const cplx *)x, *ps = (const cplx *)psi_sec, *ss = (const cplx *)sumsec;
    const cplx *fo = (const cplx *)pl->fo;
    cplx *yc = (cplx *)y;
    int64_t P = pl->P, S = P - 1, M = pl->M;
    #pragma omp parallel for schedule(static)
    for (int64_t mu = 0; mu < M; ++mu)
        for (int64_t j = 0; j < P; ++j) {
            cplx v = xc[mu * P];
            for (int64_t sc = 1; sc <= S; ++sc)
                v += fo[j * S + (sc - 1)] * (ps[mu * P + sc] - ss[sc]);
            yc[j * M + mu] = v;
        }
}

/* Whole path: FMM, combine, then a length-M FFT per block (src/fmmfft.c:1080).
 * dir < 0 computes the unnormalised inverse as conj(F(conj(x))); the
 * reference's own dir < 0 branch conjugates before its FFT
 * (src/fmmfft.c:976-977 vs :1080) and is not reproduced. */
void orc_fmmfft(const orc_plan *pl, const double *x_in, double *X, int dir) {
    int64_t N = pl->N, M = pl->M, P = pl->P;
    double *x = malloc(sizeof(double) * 2 * N);
    memcpy(x, x_in, sizeof(double) * 2 * N);
    if (dir < 0) for (int64_t i = 0; i < N; ++i) x[2 * i + 1] = -x[2 * i + 1];
    int64_t tl = orc_tree_len(pl);
    double *phi = malloc(sizeof(double) * 2 * tl);
    double *psi = malloc(sizeof(double) * 2 * tl);
    double *ps = malloc(sizeof(double) * 2 * N);
    double *ss = malloc(sizeof(double) * 2 * P);
    orc_section_sums(pl, x, ss);
    orc_upward(pl, x, phi);
    orc_downward(pl, phi, psi);
    orc_evaluate(pl, x, psi, ps);
    orc_combine(pl, x, ps, ss, X);
    #pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < P; ++j) orc_fft64(M, X + 2 * j * M, -1);
    if (dir < 0) for (int64_t i = 0; i < N; ++i) X[2 * i + 1] = -X[2 * i + 1];
    free(x); free(phi); free(psi); free(ps); free(ss);
}

/* ---------------------------------------------------------------------------
 * FP64 FFT and DFT (the oracle's stand-in for FFTW, which is not installed).
 * ------------------------------------------------------------------------- */

void orc_fft64(int64_t n, double *data, int sign) {
    cplx *a = (cplx *)data;
    if (n <= 1) return;
    for (int64_t i = 1, j = 0; i < n; ++i) {          /* bit reversal */
        int64_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { cplx t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    cplx *tw = malloc(sizeof(cplx) * (n / 2));
    for (int64_t k = 0; k < n / 2; ++k) {
        long double ang = (long double)sign * 2.0L * 3.141592653589793238462643383279502884L *
                          (long double)k / (long double)n;
        tw[k] = (double)cosl(ang) + I * (double)sinl(ang);
    }
    for (int64_t len = 2; len <= n; len <<= 1) {
        int64_t half = len >> 1, step = n / len;
        for (int64_t i = 0; i < n; i += len)
            for (int64_t k = 0; k < half; ++k) {
                cplx u = a[i + k], v = a[i + k + half] * tw[k * step];
                a[i + k] = u + v;
                a[i + k + half] = u - v;
            }
    }
    free(tw);
}

void orc_dft_direct(int64_t n, const double *in, double *out, int sign) {
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int64_t k = 0; k < n; ++k) {
        long double re = 0, im = 0;
        for (int64_t m = 0; m < n; ++m) {
            int64_t e = (k * m) % n;
            long double ang = (long double)sign * 2.0L * pi * (long double)e / (long double)n;
            long double c = cosl(ang), s = sinl(ang);
            re += in[2 * m] * c - in[2 * m + 1] * s;
            im += in[2 * m] * s + in[2 * m + 1] * c;
        }
        out[2 * k] = (double)re;
        out[2 * k + 1] = (double)im;
    }
}

void orc_cot_direct(int64_t M, int64_t P, int64_t sc, const double *x, int64_t stride, double *out) {
    const long double pi = 3.141592653589793238462643383279502884L;
    long double alpha = (long double)sc / (long double)P;
    #pragma omp parallel for schedule(static)
    for (int64_t mu = 0; mu < M; ++mu) {
        long double re = 0, im = 0;
        for (int64_t nu = 0; nu < M; ++nu) {
            long double d = (long double)(nu - mu) + alpha;
            long double k = -1.0L / tanl(pi * d / (long double)M) / (long double)M;
            re += k * x[2 * nu * stride];
            im += k * x[2 * nu * stride + 1];
        }
        out[2 * mu] = (double)re;
        out[2 * mu + 1] = (double)im;
    }
}

/* ---------------------------------------------------------------------------
 * Metrics and inputs.
 * ------------------------------------------------------------------------- */

double orc_rel_l2(int64_t n, const double *x, const double *y) {
    long double num = 0, den = 0;
    for (int64_t i = 0; i < 2 * n; ++i) {
        long double d = (long double)x[i] - y[i];
        num += d * d;
        den += (long double)y[i] * y[i];
    }
    return den > 0 ? (double)sqrtl(num / den) : (double)sqrtl(num);
}

double orc_max_abs(int64_t n, const double *x, const double *y) {
    double m = 0;
    for (int64_t i = 0; i < n; ++i) {
        double d = hypot(x[2 * i] - y[2 * i], x[2 * i + 1] - y[2 * i + 1]);
        if (d > m) m = d;
    }
    return m;
}

/* src/fmmfft_error.c:1097-1108 */
double orc_max_rel_pct(int64_t n, const double *x, const double *y) {
    double m = 0;
    for (int64_t i = 0; i < n; ++i) {
        double e = hypot(x[2 * i] - y[2 * i], x[2 * i + 1] - y[2 * i + 1]) /
                   hypot(y[2 * i], y[2 * i + 1]) * 100;
        if (i == 0 || e > m) m = e;
    }
    return m;
}

static uint64_t splitmix64(uint64_t *st) {
    uint64_t z = (*st += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double u01(uint64_t *st) { return (double)(splitmix64(st) >> 11) * 0x1.0p-53; }

void orc_fill_uniform(int64_t n, uint64_t seed, double lo, double hi, double *out) {
    uint64_t st = seed;
    for (int64_t i = 0; i < 2 * n; ++i) out[i] = lo + (hi - lo) * u01(&st);
}

void orc_fill_normal(int64_t n, uint64_t seed, double sigma, double *out) {
    uint64_t st = seed;
    for (int64_t i = 0; i < n; ++i) {
        double u1 = u01(&st), u2 = u01(&st);
        double r = sqrt(-2.0 * log1p(-u1));
        out[2 * i] = sigma * r * cos(2 * ORC_PI * u2);
        out[2 * i + 1] = sigma * r * sin(2 * ORC_PI * u2);
    }
}

/* src/fmmfft.c:1228-1246: x_n = n + 1 */
void orc_fill_ramp(int64_t n, double *out) {
    for (int64_t i = 0; i < n; ++i) { out[2 * i] = (double)(i + 1); out[2 * i + 1] = 0; }
}
