// Length-M local FFTs of the FMM-FFT (the reference's fftw_execute,
// src/fmmfft.c:1080) on gfx950.
//
// M <= 4096: one pass, one LDS-resident FFT per block row.  Larger M: the
// four-step split M = M1 * M2, mu = M2*m1 + m2, k = k1 + M1*k2:
//   pass 1: Z[k1][m2] = w_M^{m2 k1} * sum_m1 y[M2 m1 + m2] w_M1^{m1 k1}
//   pass 2: X[k1 + M1 k2] = sum_m2 Z[k1][m2] w_M2^{m2 k2}
// Each workgroup holds C sub-FFTs of length L in LDS (C consecutive columns,
// so every global access is C contiguous elements), runs an in-place radix-2
// DIT on bit-reversed input and writes the C results.  Pass-1 twiddles are
// formed and applied in FP64 from two short tables (w_M^e = hi[e>>r] lo[e&R-1]).
#pragma once
#include "common.hpp"

namespace fmmfft {

struct FftPass {
    int L, logL, C;       // sub-FFT length, log2, sub-FFTs per workgroup
    int64_t F;            // sub-FFTs per length-M block
    int64_t in_fs, in_es; // element e of sub-FFT f at k2*M + f*in_fs + e*in_es
    int64_t out_fs, out_ks;
    int twiddle;          // 1: multiply output k of sub-FFT f by w_M^{f k}
    int conj_out;         // conjugate the output (inverse transform)
    int logR;             // split of the pass-1 twiddle exponent
};

template <typename T>
__global__ __launch_bounds__(256) void k_fft(Geom g, FftPass fp, const cx<T> *__restrict__ in,
                                             cx<T> *__restrict__ out, int64_t in_stride, int64_t out_stride,
                                             const T *__restrict__ tw,  // [L/2] (re,im) w_L^k
                                             const c64 *__restrict__ tlo, const c64 *__restrict__ thi) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int L = fp.L, C = fp.C, logL = fp.logL;
    const int ls = L + 1;  // padded row
    cx<T> *a = reinterpret_cast<cx<T> *>(smem);  // [C][L+1]
    const int sig = blockIdx.z;
    in += (int64_t)sig * in_stride;
    out += (int64_t)sig * out_stride;
    const int64_t fg0 = (int64_t)blockIdx.x * C;  // first global sub-FFT
    const int64_t k2 = fg0 / fp.F;
    const int64_t f0 = fg0 - k2 * fp.F;
    const int64_t base = k2 * g.M;
    // when F == 1 the C sub-FFTs are C consecutive blocks
    const int64_t fstep_in = fp.F == 1 ? g.M : fp.in_fs;
    const int64_t fstep_out = fp.F == 1 ? g.M : fp.out_fs;

    const int n = C * L;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        int f, e;
        if (fp.in_es == 1) { e = i % L; f = i / L; } else { f = i % C; e = i / C; }
        const cx<T> v = in[base + (f0 + f) * fstep_in + (int64_t)e * fp.in_es];
        const int br = (int)(__builtin_bitreverse32((unsigned)e) >> (32 - logL));
        a[f * ls + br] = v;
    }
    __syncthreads();
    const cx<T> *twc = reinterpret_cast<const cx<T> *>(tw);
    for (int h = 1; h < L; h <<= 1) {
        const int tstep = L / (2 * h);
        for (int bf = threadIdx.x; bf < n / 2; bf += blockDim.x) {
            const int f = bf / (L / 2), t = bf - f * (L / 2);
            const int grp = t / h, k = t - grp * h;
            const int i1 = f * ls + grp * 2 * h + k, i2 = i1 + h;
            const cx<T> w = twc[k * tstep];
            const cx<T> u = a[i1];
            const cx<T> v = cmulT(a[i2], w);
            a[i1] = cx<T>{u.re + v.re, u.im + v.im};
            a[i2] = cx<T>{u.re - v.re, u.im - v.im};
        }
        __syncthreads();
    }
    const int64_t Rm = (int64_t(1) << fp.logR) - 1;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        int f, k;
        if (fp.out_ks == 1) { k = i % L; f = i / L; } else { f = i % C; k = i / C; }
        cx<T> v = a[f * ls + k];
        if (fp.twiddle) {
            const int64_t e = ((f0 + f) * (int64_t)k) & (g.M - 1);
            const c64 w = cmul(thi[e >> fp.logR], tlo[e & Rm]);
            const c64 r = cmul(c64{(double)v.re, (double)v.im}, w);
            v = cx<T>{(T)r.re, (T)r.im};
        }
        if (fp.conj_out) v.im = -v.im;
        out[base + (f0 + f) * fstep_out + (int64_t)k * fp.out_ks] = v;
    }
}

}  // namespace fmmfft
