// Length-M local FFTs of the FMM-FFT (the reference's fftw_execute,
// src/fmmfft.c:1080) on gfx950.
//
// M <= 1024: one pass, one LDS-resident FFT per block row.  Larger M: the
// four-step split M = M1 * M2, mu = M2*m1 + m2, k = k1 + M1*k2:
//   pass 1: Z[k1][m2] = w_M^{m2 k1} * sum_m1 y[M2 m1 + m2] w_M1^{m1 k1}
//   pass 2: X[k1 + M1 k2] = sum_m2 Z[k1][m2] w_M2^{m2 k2}
// A workgroup holds C sub-FFTs of length L in LDS; they are C consecutive
// columns (or rows), so every global access moves C (or L) contiguous
// elements.  Inside LDS the transform is a Stockham autosort FFT with radix
// 16/8/4 butterflies held in registers (three LDS round trips for L = 1024).
// Pass-1 twiddles are formed and applied in FP64 from two short tables
// (w_M^e = hi[e >> r] * lo[e & (R-1)]).
#pragma once
#include "common.hpp"

namespace fmmfft {

struct FftPass {
    int64_t F;             // sub-FFTs per length-M block
    int64_t in_fs, in_es;  // element e of sub-FFT f at k2*M + f*in_fs + e*in_es;
    int64_t in_bs;         // column-tiled input: f = f0 + j (f0 the tile's first,
    int in_ltw;            // a multiple of its C), f0 = q 2^in_ltw + r, at
                           // k2*M + q*in_bs + (r + j)*in_fs + e*in_es
                           // (in_ltw = 30: q = 0, the plain form)
    int64_t out_fs, out_ks;
    int64_t out_bs;        // column-tiled output (out_tile > 0): sub-FFT f at
    int out_tile;          // (f >> out_tile)*out_bs + (f & (2^out_tile - 1))*out_fs
    int twiddle;           // 1: multiply output k of sub-FFT f by w_M^{g(f) k},
    int tw_lf, tw_shift;   // g(f) = (f mod 2^tw_lf) * 2^tw_shift (tw_lf = 0: f)
    int conj_out;          // conjugate the output (inverse transform)
    int logR;              // split of the pass-1 twiddle exponent
};

template <typename T>
__device__ inline cx<T> cadd(cx<T> a, cx<T> b) { return cx<T>{a.re + b.re, a.im + b.im}; }
template <typename T>
__device__ inline cx<T> csub(cx<T> a, cx<T> b) { return cx<T>{a.re - b.re, a.im - b.im}; }
// multiply by -i
template <typename T>
__device__ inline cx<T> cmi(cx<T> a) { return cx<T>{a.im, -a.re}; }

// Forward radix-R DFTs in registers: y_q = sum_r x_r e^{-2 pi i r q / R}.
template <typename T>
__device__ inline void dft2(cx<T> &a, cx<T> &b) {
    cx<T> t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

template <typename T>
__device__ inline void dft4(cx<T> *v) {
    cx<T> s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    cx<T> s13 = cadd(v[1], v[3]), d13 = cmi(csub(v[1], v[3]));
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
}

template <typename T>
__device__ inline void dft8(cx<T> *v) {
    const T h = T(0.70710678118654752440084436210484903928);
    cx<T> e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    dft4(e);
    dft4(o);
    // w8^1 = (1 - i) h, w8^2 = -i, w8^3 = -(1 + i) h
    cx<T> o1{(o[1].re + o[1].im) * h, (o[1].im - o[1].re) * h};
    cx<T> o2 = cmi(o[2]);
    cx<T> o3{(o[3].im - o[3].re) * h, -(o[3].re + o[3].im) * h};
    v[0] = cadd(e[0], o[0]); v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);   v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);   v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);   v[7] = csub(e[3], o3);
}

template <typename T>
__device__ inline void dft16(cx<T> *v) {
    // w16^q = cos(pi q / 8) - i sin(pi q / 8)
    const T c1 = T(0.92387953251128675612818318939678828682), s1 = T(0.38268343236508977172845998403039886676);
    const T h = T(0.70710678118654752440084436210484903928);
    cx<T> e[8], o[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        e[i] = v[2 * i];
        o[i] = v[2 * i + 1];
    }
    dft8(e);
    dft8(o);
    const cx<T> w[8] = {{T(1), T(0)}, {c1, -s1}, {h, -h}, {s1, -c1}, {T(0), T(-1)}, {-s1, -c1}, {-h, -h}, {-c1, -s1}};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        cx<T> t = q == 0 ? o[0] : cmulT(o[q], w[q]);
        v[q] = cadd(e[q], t);
        v[q + 8] = csub(e[q], t);
    }
}

template <typename T, int R>
__device__ inline void dftR(cx<T> *v) {
    if constexpr (R == 2) dft2(v[0], v[1]);
    else if constexpr (R == 4) dft4(v);
    else if constexpr (R == 8) dft8(v);
    else dft16(v);
}

// Radix of the next Stockham stage for the remaining length rem.  FP64
// stays at radix <= 8: a radix-16 butterfly with its 15 twiddles would take
// ~190 registers per lane in FP64 and spill; the extra LDS round trip is
// cheap next to the HBM time of a tile.
template <typename T>
constexpr int pick_radix(int rem) {
    if (sizeof(T) == 8) return rem == 2 ? 2 : rem == 4 ? 4 : rem == 16 ? 4 : 8;
    return rem == 2 ? 2 : rem == 4 ? 4 : rem == 8 ? 8 : rem == 32 ? 8 : rem == 512 ? 8 : 16;
}

// LDS index of element e of a row: one pad slot per 2^PS elements breaks the
// bank aliasing of the radix-2^PS stride (16 for the local FFTs; 8 for the
// combine's radix-8 FP64 stages, whose first stage writes at a stride of
// 8 elements of 16 bytes)
template <int PS>
__device__ inline int lpadp(int e) { return e + (e >> PS); }
__device__ inline int lpad(int e) { return lpadp<4>(e); }

// One Stockham stage (current sub-transform size Ns, radix R) over C rows,
// NT threads.
template <typename T, int L, int C, int Ns, int R, int NT, int PS = 4>
__device__ inline void stockham_stage(cx<T> *a, int ls, const cx<T> *__restrict__ twL) {
    constexpr int NB = L / R;
    constexpr int TOT = C * NB;
    constexpr int BPT = (TOT + NT - 1) / NT;
    cx<T> v[BPT][R];
#pragma unroll
    for (int bq = 0; bq < BPT; ++bq) {
        const int w = threadIdx.x + bq * NT;
        if (TOT % NT == 0 || w < TOT) {
            const int f = w / NB, j = w % NB;
            const cx<T> *row = a + f * ls;
            const int k = j & (Ns - 1);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                cx<T> x = row[lpadp<PS>(j + r * NB)];
                if (Ns > 1 && r > 0) x = cmulT(x, twL[r * k * (L / (Ns * R))]);
                v[bq][r] = x;
            }
            dftR<T, R>(v[bq]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int bq = 0; bq < BPT; ++bq) {
        const int w = threadIdx.x + bq * NT;
        if (TOT % NT == 0 || w < TOT) {
            const int f = w / NB, j = w % NB;
            cx<T> *row = a + f * ls;
            const int k = j & (Ns - 1);
            const int d = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) row[lpadp<PS>(d + r * Ns)] = v[bq][r];
        }
    }
    __syncthreads();
}

template <typename T, int L, int C, int Ns, int NT, int PS = 4>
__device__ inline void stockham(cx<T> *a, int ls, const cx<T> *__restrict__ twL) {
    if constexpr (Ns < L) {
        constexpr int R = pick_radix<T>(L / Ns);
        stockham_stage<T, L, C, Ns, R, NT, PS>(a, ls, twL);
        stockham<T, L, C, Ns * R, NT, PS>(a, ls, twL);
    }
}

// Row stride (in elements) of a tile of C rows of length L holding elements
// of `esize` bytes.  Inside a row one pad slot per 2^PS elements breaks
// the radix-2^PS stride (16 for complex64, 8 for complex128 whose stages
// are radix <= 8: its passes 0.463/0.552 -> 0.455/0.547 ms at M = 2^20);
// across rows the stride is chosen so that a transposed
// access (consecutive lanes step through C' = min(C, Q) rows, then along
// the row) hits Q distinct bank groups per LDS cycle, Q = 128 / esize lanes:
// stride = Q / C' (mod Q).
template <int L, int C, int ESIZE, int PS = 4>
constexpr int tile_stride() {
    constexpr int Q = 128 / ESIZE;
    constexpr int CQ = C < Q ? C : Q;
    constexpr int base = L + (L >> PS);
    constexpr int want = Q / CQ;
    return base + (((want - base) % Q) + Q) % Q;
}

// Sub-FFTs per workgroup: 8192 (complex64) or 4096 (complex128) elements,
// at most 32; about 70 KB of LDS.
template <typename T>
constexpr int fft_cols(int L) {
    const int elems = sizeof(T) == 4 ? 8192 : 4096;
    return elems / L > 32 ? 32 : elems / L;
}
template <typename T, int L>
constexpr int fft_cols() { return fft_cols<T>(L); }

// Workgroup index remap: the round-robin dispatch puts consecutive workgroup
// ids on different XCDs (each with its own L2); remap so that each XCD runs a
// contiguous range of logical ids, i.e. neighbouring column chunks that
// share 128-byte lines are served by one L2.
__device__ inline int64_t xcd_remap(int64_t bid, int64_t nwg) {
    return (nwg % 8 == 0) ? (bid % 8) * (nwg / 8) + bid / 8 : bid;
}

// C sub-FFTs of length L per tile; twL = e^{-2 pi i k / L}, k < L.
// Persistent: a workgroup loops over tiles (signal = blockIdx.z) and issues
// the global loads of its next tile into registers before transforming the
// current one, so HBM latency hides behind the LDS work.  Each lane's output
// elements share one sub-FFT f and step k by NT/C, so the pass-1 twiddle
// w_M^{f k} is a per-lane FP64 recurrence started from two table lookups.
// Offsets inside one signal fit 32 bits (N <= 2^30), which keeps the
// per-element address arithmetic to one register.
constexpr int kFftThreads = 512;

// Access forms of a pass: element rows / output rows contiguous or strided.
enum FftForm { FFT_SINGLE = 0, FFT_PASS1 = 1, FFT_PASS2 = 2 };

// PERSIST = false: one tile per workgroup (grid = tiles), no prefetch.
// V2 (complex64 pass 1, 16-byte aligned input with even column, element and
// block strides): a lane loads two neighbouring columns of an element as one
// 16-byte access, halving the pass's load instructions (N = 2^26: pass 1
// 0.252 -> 0.242 ms; its texture-address unit was the one held up).
// S2 (complex64 last pass, 16-byte aligned output with even strides): a lane
// stores two neighbouring sub-FFTs' outputs of one k as one 16-byte access.
template <typename T, int L, int FORM, int C = fft_cols<T, L>(), bool PERSIST = true, bool TAB = false,
          bool V2 = false, bool S2 = false>
__global__ __launch_bounds__(kFftThreads) void k_fft(Geom g, FftPass fp, SigAddr<const cx<T>> in,
                                             SigAddr<cx<T>> out,
                                             const cx<T> *__restrict__ twL, const c64 *__restrict__ tlo,
                                             const c64 *__restrict__ thi, int nsig) {
    constexpr int PSX = sizeof(T) == 8 ? 3 : 4;  // FP64 stages are radix <= 8
    constexpr int ls = tile_stride<L, C, (int)sizeof(cx<T>), PSX>();
    constexpr int n = C * L;
    constexpr int NT = kFftThreads;
    constexpr int PER = n / NT;
    static_assert(n % NT == 0 && NT % C == 0, "tile shape");
    __shared__ cx<T> a[C * ls];
    const int M = (int)g.M, F = (int)fp.F;
    const int nf = F * (int)g.P;  // sub-FFTs per signal
    const int ntiles = (nf + C - 1) / C;
    const int ntot = nsig * ntiles;  // tiles of all signals: tile = signal * ntiles + t
    const int nwg = gridDim.x;
    const int xo = (int)xcd_remap(blockIdx.x, nwg);
    // when F == 1 the C sub-FFTs of a tile are C consecutive blocks
    const int fstep_in = F == 1 ? M : (int)fp.in_fs;
    const int fblk_in = (int)fp.in_bs, in_ltw = fp.in_ltw;
    const int fstep_out = F == 1 ? M : (int)fp.out_fs;
    const int in_es = (int)fp.in_es, out_ks = (int)fp.out_ks;
    // single pass: e and k run along rows; pass 1: both strided (columns of
    // [M1][M2]); pass 2: e along rows, k strided
    constexpr bool erow = FORM != FFT_PASS1, krow = FORM == FFT_SINGLE;
    // complex64: the last pass (single, pass 2 or 3) loads and stores with
    // the nontemporal hint, so the transform's output does not sit dirty in
    // the last-level cache for the next kernel to write back (N = 2^26: the
    // next transform's S2M 0.185 -> 0.147 ms, the whole transform 1.554 ->
    // 1.490 ms).  complex128 keeps cached accesses: its pass 2 (one tile per
    // workgroup, 64-byte output pieces) measured 0.54 -> 0.62 ms with them.
    constexpr bool NT_LAST = sizeof(T) == 4;
    const int Rm = (1 << fp.logR) - 1;

    // re/im in separate scalar arrays and unconditional loads (a ragged last
    // tile re-reads its last valid sub-FFT): the prefetch stays in registers.
    T vr[PER], vi[PER];
    static_assert(!V2 || (FORM == FFT_PASS1 && sizeof(T) == 4 && C % 2 == 0), "two-column loads: complex64 pass 1");
    static_assert(!S2 || (FORM == FFT_PASS2 && sizeof(T) == 4 && C % 2 == 0), "two-column stores: complex64 pass 2");
    auto prefetch = [&](int tt) __attribute__((always_inline)) {
        const int sig = tt / ntiles, t = tt - sig * ntiles;
        const cx<T> *in_s = in.template at<TAB>(sig);
        const int fg0 = t * C;
        const int cl = min(C, nf - fg0) - 1;
        const int k2 = fg0 / F, f0 = fg0 - k2 * F, base = k2 * M;
        if constexpr (V2) {
            // columns f, f + 1 (f even; a tile holds an even number of
            // columns, so f <= cl - 1) of element e
#pragma unroll
            for (int u = 0; u < PER / 2; ++u) {
                const int i = threadIdx.x + u * NT;
                const int f = min(2 * (i % (C / 2)), cl - 1), e = i / (C / 2);
                const int fo = (f0 >> in_ltw) * fblk_in + ((f0 & ((1 << in_ltw) - 1)) + f) * fstep_in;
                const float4 v = *reinterpret_cast<const float4 *>(in_s + base + fo + e * in_es);
                vr[2 * u] = v.x;
                vi[2 * u] = v.y;
                vr[2 * u + 1] = v.z;
                vi[2 * u + 1] = v.w;
            }
            return;
        }
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int i = threadIdx.x + u * NT;
            const int f = erow ? i / L : i % C, e = erow ? i % L : i / C;
            const int fo = (f0 >> in_ltw) * fblk_in + ((f0 & ((1 << in_ltw) - 1)) + min(f, cl)) * fstep_in;
            const cx<T> x = (NT_LAST && FORM != FFT_PASS1) ? ld_nt(in_s + base + fo + e * in_es)
                                                          : in_s[base + fo + e * in_es];
            vr[u] = x.re;
            vi[u] = x.im;
        }
    };

    int tt = xo;
    if (tt < ntot) prefetch(tt);
    while (tt < ntot) {
        __syncthreads();  // the previous tile's outputs have left LDS
        if constexpr (V2) {
#pragma unroll
            for (int u = 0; u < PER / 2; ++u) {
                const int i = threadIdx.x + u * NT;
                const int f = 2 * (i % (C / 2)), e = i / (C / 2);
                a[f * ls + lpadp<PSX>(e)] = cx<T>{vr[2 * u], vi[2 * u]};
                a[(f + 1) * ls + lpadp<PSX>(e)] = cx<T>{vr[2 * u + 1], vi[2 * u + 1]};
            }
        } else {
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                const int i = threadIdx.x + u * NT;
                const int f = erow ? i / L : i % C, e = erow ? i % L : i / C;
                a[f * ls + lpadp<PSX>(e)] = cx<T>{vr[u], vi[u]};
            }
        }
        __syncthreads();
        const int sig = tt / ntiles, t = tt - sig * ntiles;
        cx<T> *out_s = out.template at<TAB>(sig);
        const int fg0 = t * C;
        const int cv = min(C, nf - fg0);
        const int k2 = fg0 / F, f0 = fg0 - k2 * F, base = k2 * M;
        // pass-1 twiddle factors of this tile, loaded before the next tile's
        // prefetch so that their use waits for them alone
        // (lane's f = tid % C fixed, k = tid / C + u * (NT / C)).  complex128
        // keeps the loads and the run-time twiddle test in every form:
        // compiling them out of pass 2 changed the schedule of its output
        // loop and measured 0.54 -> 0.61 ms at M = 2^20.
        constexpr bool TWF = FORM == FFT_PASS1 || sizeof(T) == 8;
        c64 th0{1.0, 0.0}, tl0{1.0, 0.0}, ths{1.0, 0.0}, tls{1.0, 0.0};
        if constexpr (TWF) {
            int fa = f0 + (int)threadIdx.x % C;
            if (fp.tw_lf) fa &= (1 << fp.tw_lf) - 1;
            const int64_t fs = (int64_t)fa << fp.tw_shift;
            const int e0 = (int)((fs * (threadIdx.x / C)) & (M - 1));
            const int es = (int)((fs * (NT / C)) & (M - 1));
            th0 = thi[e0 >> fp.logR];
            tl0 = tlo[e0 & Rm];
            ths = thi[es >> fp.logR];
            tls = tlo[es & Rm];
        }
        // past the last tile the prefetch re-reads the current one (still in
        // L2, and the prefetch stays unconditional, which keeps the
        // counter waits exact)
        const int tn = tt + nwg;
        if constexpr (PERSIST) prefetch(tn < ntot ? tn : tt);
        stockham<T, L, C, 1, NT, PSX>(a, ls, twL);

        if constexpr (S2) {
            // (no pass-2 twiddle in complex64; out_fs = 1, so sub-FFTs f and
            // f + 1 are neighbours; a tile holds an even number of them)
            typedef float f4v __attribute__((ext_vector_type(4)));
#pragma unroll
            for (int u = 0; u < PER / 2; ++u) {
                const int i = threadIdx.x + u * NT;
                const int f = 2 * (i % (C / 2)), k = i / (C / 2);
                const cx<T> v0 = a[f * ls + lpadp<PSX>(k)], v1 = a[(f + 1) * ls + lpadp<PSX>(k)];
                const float sg = fp.conj_out ? -1.f : 1.f;
                if (f < cv)
                    __builtin_nontemporal_store(f4v{v0.re, sg * v0.im, v1.re, sg * v1.im},
                                                reinterpret_cast<f4v *>(out_s + base + (f0 + f) + k * out_ks));
            }
            tt = tn;
            continue;
        }
        c64 w{1.0, 0.0}, ws{1.0, 0.0};
        if (TWF && fp.twiddle) {
            w = cmul(th0, tl0);
            ws = cmul(ths, tls);
        }
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int i = threadIdx.x + u * NT;
            const int f = krow ? i / L : i % C, k = krow ? i % L : i / C;
            cx<T> val = a[f * ls + lpadp<PSX>(k)];
            if (TWF && fp.twiddle) {
                // FP64 recurrence, applied in the element's precision
                // (complex64: an FP32 product with the rounded factor)
                if constexpr (sizeof(T) == 4) {
                    val = cmulT(val, cx<T>{(T)w.re, (T)w.im});
                } else {
                    const c64 r = cmul(c64{(double)val.re, (double)val.im}, w);
                    val = cx<T>{(T)r.re, (T)r.im};
                }
                w = cmul(w, ws);
            }
            if (fp.conj_out) val.im = -val.im;
            const int fg = f0 + f;
            const int fo = fp.out_tile ? (fg >> fp.out_tile) * (int)fp.out_bs + (fg & ((1 << fp.out_tile) - 1)) * fstep_out
                                       : fg * fstep_out;
            if (f < cv) {
                if (NT_LAST && FORM != FFT_PASS1) st_nt(out_s + base + fo + k * out_ks, val);
                else out_s[base + fo + k * out_ks] = val;
            }
        }
        tt = tn;
    }
}

}  // namespace fmmfft
