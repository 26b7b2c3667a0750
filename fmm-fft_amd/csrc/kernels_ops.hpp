// Plan-time operator construction on the device, in FP64 (gfx950).
//
// The FMM's translation operators are rational interpolation matrices with
// prescribed poles (the reference builds them on the host in
// src/fmmfft.c:81-172, 186-197, 227-259 and assembles them in mftii,
// :261-332).  Here every matrix (or evaluation row) is one workgroup: the
// wave computes the Chebyshev nodes, the poles and the evaluation points in
// LDS, then one lane per entry forms
//   mat[k][j] = w_j / (x_k - pole_j) * prod_{l != j} (x_k - c_l) / (x_k - pole_l)
//   w_j       = (c_j - pole_j) * prod_{l != j} (c_j - pole_l) / (c_j - c_l).
// Closed-form tables (S2M, near-field kernel, combine factors, FFT twiddles)
// are one lane per entry.
#pragma once
#include "common.hpp"

namespace fmmfft {

namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr int kMaxTerms = 32;

// Multipole-side poles (src/fmmfft.c:82-87).
__device__ void poles_multipole(double a, int terms, double *pole, int j) {
    double ac = 3.0 * tan(a);
    double th = -1.0 * cos(((double)(2 * (j + 1)) - 2) / (2 * (terms - 1)) * kPi) * (2 * a);
    pole[j] = ac / tan(0.5 * th);
}

// Local-side poles (src/fmmfft.c:117-123).
__device__ void poles_local(double a, int terms, double *pole, int j) {
    double ac = 1.0 / tan(a);
    double th = -cos(((double)(2 * (j + 1) - 2)) / (2 * (terms - 1)) * kPi) * (kPi - 6 * a) + kPi;
    pole[j] = ac * tan(0.5 * th);
}

}  // namespace

// Job layout of one launch (blockIdx.x):
//   [0, 8(n-2))          level operators, slot L = job / 8, op = job % 8
//   8(n-2)               top flip (level 2, shift 2)
//   next S*b             evalm rows (sc, j)
//   next P/2             evalmh rows
__global__ __launch_bounds__(64) void k_build_interp(Geom g, double *ops, double *topflp,
                                                      double *evalm, double *evalmh) {
    __shared__ double cheb[kMaxTerms], pole[kMaxTerms], xs[kMaxTerms], w[kMaxTerms];
    const int terms = g.terms;
    const int lane = threadIdx.x;
    int job = blockIdx.x;

    if (lane < terms) cheb[lane] = -1 * cos(((double)(2 * (lane + 1)) - 1) / (2 * terms) * kPi);
    __syncthreads();

    double *dst;
    int nx = terms;
    double scale = 1.0;
    const int nlev = 8 * (g.n - 2);
    if (job <= nlev) {
        int lev, op;
        if (job < nlev) {
            lev = job / 8 + 3;  // flips / M2M of level slot L act at level L+3
            op = job % 8;
            dst = ops + (size_t)job * terms * terms;
        } else {
            lev = 2;
            op = -1;  // top flip
            dst = topflp;
        }
        double a, dd;
        if (op == OP_SHFTLN || op == OP_SHFTLP) {
            // L2L into a child at level L+3 uses the reference's shftl(L+2)
            // (src/fmmfft.c:186-197, built at :288-289).
            a = kPi / pow(2.0, (double)(lev - 1));
            dd = op == OP_SHFTLN ? -1.0 : 1.0;
            double c = tan(0.25 * a), td2 = tan(0.5 * a);
            if (lane < terms) {
                xs[lane] = (cheb[lane] - dd) / (1 / c + dd * c * cheb[lane]) / td2;
                poles_local(0.5 * a, terms, pole, lane);
            }
        } else if (op == OP_SHFTFN || op == OP_SHFTFP) {
            // M2M from a child at level L+3 (src/fmmfft.c:148-159, :282-283).
            a = kPi / pow(2.0, (double)lev);
            dd = op == OP_SHFTFN ? -1.0 : 1.0;
            double ta3 = 3 * tan(a), td2 = tan(0.5 * a);
            if (lane < terms) {
                xs[lane] = 3 * td2 * (cheb[lane] + dd * ta3 * td2) / (ta3 - dd * td2 * cheb[lane]);
                poles_multipole(0.5 * a, terms, pole, lane);
            }
        } else {
            // Flips (src/fmmfft.c:227-240, :279, :284-287).
            int shft = op == OP_FLIP2N ? -2 : op == OP_FLIP2P ? 2 : op == OP_FLIP3N ? -3 : op == OP_FLIP3P ? 3 : 2;
            a = kPi / pow(2.0, (double)lev);
            double c = tan(0.5 * a), td = tan((double)shft * a);
            if (lane < terms) {
                xs[lane] = 3 * c * (1 - td * c * cheb[lane]) / (td + c * cheb[lane]);
                poles_multipole(0.5 * a, terms, pole, lane);
            }
        }
    } else {
        // L2T rows (src/fmmfft.c:161-172, 292-303): target at box
        // coordinate ffr of a leaf with half-width pi/(2s), scaled by 1/M.
        job -= nlev + 1;
        double ffr;
        const int b = g.b, P = g.P;
        if (job < g.S * b) {
            int sc = job / b + 1, j = job % b + 1;
            ffr = ((double)(2 * j - 1) - ((double)(2 * sc)) / (double)P) / (double)b - 1;
            dst = evalm + (size_t)job * terms;
        } else {
            int e = job - g.S * b;  // sc = P/2 + e
            int sc = P / 2 + e;
            ffr = (((double)2 * b + 1) - ((double)2 * sc) / (double)P) / (double)b - 1;
            dst = evalmh + (size_t)e * terms;
        }
        double a = (kPi / (double)g.s) * 0.5;
        nx = 1;
        scale = (double)g.M;
        if (lane == 0) xs[0] = tan(ffr * a) / tan(a);
        if (lane < terms) poles_local(a, terms, pole, lane);
    }
    __syncthreads();

    if (lane < terms) {
        double v = cheb[lane] - pole[lane];
        for (int l = 0; l < terms; ++l)
            if (l != lane) v *= (cheb[lane] - pole[l]) / (cheb[lane] - cheb[l]);
        w[lane] = v;
    }
    __syncthreads();

    for (int e = lane; e < nx * terms; e += 64) {
        int k = e / terms, j = e % terms;
        double xk = xs[k];
        double v = w[j] / (xk - pole[j]);
        for (int l = 0; l < terms; ++l)
            if (l != j) v *= (xk - cheb[l]) / (xk - pole[l]);
        dst[e] = nx == 1 ? v / scale : v;
    }
}

// Closed-form tables, one lane per entry:
//   cheb [terms]; initch [b][terms] (src/fmmfft.c:242-259);
//   kv [S][4b] = -1/tan(pi/M (d + sc/P))/M, d = -2b .. 2b-1 -- the
//   generator of the reference's Toeplitz tables cots / cotprv / cotsh
//   (src/fmmfft.c:305-320);
//   gfac [P] = i (1 - w_P^sc) / 2 = -e^{-i pi sc/P} sin(pi sc/P), the fo
//   factor of src/fmmfft.c:325-331 with its w_P^{sc(j-1)} part left to the
//   P-point DFT;  wP [P] = e^{-2 pi i k/P}.
__global__ __launch_bounds__(256) void k_build_closed(Geom g, double *cheb, double *initch,
                                                       double *kv, double2 *gfac, double2 *wP) {
    const int terms = g.terms, b = g.b, P = g.P;
    const int64_t n_cheb = terms, n_init = (int64_t)b * terms, n_kv = (int64_t)g.S * 4 * b;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_cheb) {
        cheb[i] = -1 * cos(((double)(2 * (i + 1)) - 1) / (2 * terms) * kPi);
        return;
    }
    i -= n_cheb;
    if (i < n_init) {
        int j = (int)(i / terms) + 1, k = (int)(i % terms);
        double ck = -1 * cos(((double)(2 * (k + 1)) - 1) / (2 * terms) * kPi);
        double a = kPi / (2.0 * (double)g.s);
        double ta3 = 3 * tan(a);
        double tg = tan(((double)(2 * j - b - 1)) / (double)b * a);
        initch[i] = (ck + ta3 * tg) / (ta3 - ck * tg);
        return;
    }
    i -= n_init;
    if (i < n_kv) {
        int sc = (int)(i / (4 * b)) + 1, d = (int)(i % (4 * b)) - 2 * b;
        double dM = (double)g.M;
        double gfrac = (double)sc / (double)P;
        kv[i] = -1 / tan(kPi / dM * ((double)d + gfrac)) / dM;
        return;
    }
    i -= n_kv;
    if (i < P) {
        int sc = (int)i;
        double sn, cs;
        sincospi((double)sc / (double)P, &sn, &cs);
        // -e^{-i pi sc/P} sin(pi sc/P)
        gfac[sc] = sc == 0 ? make_double2(0.0, 0.0) : make_double2(-cs * sn, sn * sn);
        double s2, c2;
        sincospi(2.0 * (double)sc / (double)P, &s2, &c2);
        wP[sc] = make_double2(c2, -s2);
    }
}

// FFT twiddles: tw[k] = e^{-2 pi i k / L} for k < L (stage twiddles of a
// sub-FFT), lo[k] = e^{-2 pi i k / M} for k < R, hi[k] = e^{-2 pi i k R / M}
// for k < M/R (their product gives any e^{-2 pi i e / M}).
__global__ __launch_bounds__(256) void k_build_twiddles(int64_t L, int64_t M, int64_t R, double2 *tw,
                                                         double2 *lo, double2 *hi) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double s, c;
    if (i < L) {
        sincospi(2.0 * (double)i / (double)L, &s, &c);
        tw[i] = make_double2(c, -s);
        return;
    }
    i -= L;
    if (R <= 0) return;  // stage twiddles only
    if (i < R) {
        sincospi(2.0 * (double)i / (double)M, &s, &c);
        lo[i] = make_double2(c, -s);
        return;
    }
    i -= R;
    if (i < M / R) {
        sincospi(2.0 * (double)(i * R) / (double)M, &s, &c);
        hi[i] = make_double2(c, -s);
    }
}

// L2T rows in MFMA fragment order for the matrix-core leaf kernel:
// evfrag[sc-1][rt][s][lane] = evalm[sc-1][16 rt + lane % 16][4 s + lane / 16]
// (zero past the last term), so that one K step of one 16-row tile is one
// coalesced 256-byte load.
__global__ __launch_bounds__(256) void k_build_evfrag(Geom g, const double *evalm, double *evfrag, int nrt,
                                                       int kt) {
    const int64_t n = (int64_t)g.S * nrt * kt * 64;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int lane = (int)(e & 63);
    int64_t r = e >> 6;
    const int s = (int)(r % kt);
    r /= kt;
    const int rt = (int)(r % nrt);
    const int64_t sc1 = r / nrt;
    const int k = 4 * s + lane / 16, j = 16 * rt + lane % 16;
    evfrag[e] = k < g.terms ? evalm[(sc1 * g.b + j) * g.terms + k] : 0.0;
}

// Tree operators in matrix-core fragment order for k_tree_mma: op j (slot
// L = j / NOPS, operator j % NOPS; j = (n-2) NOPS is topflp), K step s,
// output tile ko, lane l: A[k = 16 ko + l%16][kk = 4 s + l/16], zero past
// terms.  tp = terms padded to 16 or 32; one operator is tp*tp entries.
__global__ __launch_bounds__(256) void k_build_opfrag(Geom g, const double *ops, const double *topflp,
                                                      double *opfrag, int tp) {
    const int nop = (g.n - 2) * NOPS + 1;
    const int per = tp * tp, ko_n = tp / 16;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nop * per) return;
    const int j = (int)(e / per), r = (int)(e % per);
    const int lane = r & 63, f = r >> 6;  // f = s * ko_n + ko
    const int s = f / ko_n, ko = f % ko_n;
    const int k = 16 * ko + (lane & 15), kk = 4 * s + (lane >> 4);
    const int T = g.terms;
    const double *op = j < nop - 1 ? ops + (size_t)j * T * T : topflp;
    opfrag[e] = (k < T && kk < T) ? op[k * T + kk] : 0.0;
}

__global__ __launch_bounds__(256) void k_to_f32(const double *src, float *dst, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (float)src[i];
}

// The FP32 residual of an FP64 table: dst[i] = src[i] - (float)src[i]
// rounded to FP32, so that (float)src + dst carries src to ~2^-48.
__global__ __launch_bounds__(256) void k_to_f32_lo(const double *src, float *dst, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (float)(src[i] - (double)(float)src[i]);
}

// Near-field kernel samples of the FP16 leaf kernel in MFMA fragment order.
// For section sc, fragment f (Toeplitz offset m = f - (nrt - 1)), part h
// (0 hi, 1 lo) and lane l = 16 kr + col, element i < 4 holds the FP16 part
// of 2^14 K_sc(d) with kv index 16 m + 4 kr + i - col + b (d = index - 2b):
// the A operand of v_mfma_f32_16x16x16_f16 for row tile rt and K step ks
// with m = ks - rt.  The two singular samples (d = 0 and d = -1, where
// |d + sc/P| < 1) are zero here: the kernel adds them in FP64.
// hi = fp16(v), lo = fp16(v - hi), so hi + lo = v to ~2^-22.
__global__ __launch_bounds__(256) void k_build_kfrag(Geom g, const double *kv, _Float16 *kfrag, int nrt, int nf) {
    const int64_t n = (int64_t)g.S * nf * 2 * 64 * 4;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int i = (int)(e & 3), lane = (int)((e >> 2) & 63);
    int64_t r = e >> 8;
    const int h = (int)(r & 1);
    r >>= 1;
    const int f = (int)(r % nf);
    const int64_t sc1 = r / nf;
    const int b = g.b, m = f - (nrt - 1);
    const int idx = 16 * m + 4 * (lane >> 4) + i - (lane & 15) + b;
    double v = 0.0;
    if (idx >= 0 && idx < 4 * b && idx != 2 * b && idx != 2 * b - 1) v = ldexp(kv[sc1 * 4 * b + idx], 14);
    const _Float16 hi = (_Float16)(float)v;
    kfrag[e] = h == 0 ? hi : (_Float16)(float)(v - (double)(float)hi);
}

}  // namespace fmmfft
