// FMM passes of the FMM-FFT on gfx950: S2M with section sums, the upward
// (M2M) and downward (top flip, M2L, L2L) tree passes, the leaf evaluation
// (L2T + near field + half-shift rule) and the section combine.
//
// Layouts (one signal; batched launches add blockIdx.z * stride):
//   x     [M][P]  complex, natural order: section sc is column sc
//                 (the reference's polyphase permute, src/fmmfft.c:419-439,
//                 is only an index map here)
//   tree  levels 2..n, level l has 2^l nodes of [P-1][2][terms] reals
//                 (re and im planes of each section's expansion)
//   psec  [M][P]  complex: column sc >= 1 is the cot sum of section sc,
//                 column 0 is a copy of x_0
//   y     [P][M]  complex: block k2 feeds the k2-th length-M FFT
#pragma once
#include "common.hpp"

namespace fmmfft {

// ---------------------------------------------------------------------------
// S2M + per-block section partial sums.
// The reference's initmm (src/fmmfft.c:199-211, called at :465):
//   phi_leaf[box][sc][k] = sum_j x_sc[box*b + j] * initch[j][k]
// and its per-rank section sums (src/fmmfft.c:441-450), accumulated in FP64.
// One workgroup handles `bpb` consecutive leaf boxes; a box is one
// contiguous [b][P] tile of x.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_s2m(Geom g, const cx<T> *__restrict__ x, int64_t x_stride,
                                             T *__restrict__ phi, int64_t tree_stride,
                                             c64 *__restrict__ sums_part, int64_t sums_stride,
                                             const T *__restrict__ initch, int bpb, int conj) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int P = g.P, b = g.b, terms = g.terms, S = g.S;
    cx<T> *tile = reinterpret_cast<cx<T> *>(smem);  // [b][P]
    const int sig = blockIdx.z;
    x += (int64_t)sig * x_stride;
    phi += (int64_t)sig * tree_stride + level_off(g, g.n);
    sums_part += (int64_t)sig * sums_stride + (int64_t)blockIdx.x * P;

    double sre = 0.0, sim = 0.0;  // thread t < P: running sum of section t
    const int box0 = blockIdx.x * bpb;
    const int nout = S * 2 * terms;
    for (int bx = 0; bx < bpb; ++bx) {
        const int box = box0 + bx;
        if (box >= g.s) break;
        __syncthreads();
        const cx<T> *src = x + (int64_t)box * b * P;
        for (int i = threadIdx.x; i < b * P; i += blockDim.x) {
            cx<T> v = src[i];
            if (conj) v.im = -v.im;
            tile[i] = v;
        }
        __syncthreads();
        if (threadIdx.x < P) {
            for (int j = 0; j < b; ++j) {
                cx<T> v = tile[j * P + threadIdx.x];
                sre += (double)v.re;
                sim += (double)v.im;
            }
        }
        T *dst = phi + (int64_t)box * nout;
        for (int o = threadIdx.x; o < nout; o += blockDim.x) {
            const int row = o / terms, k = o - row * terms;
            const int sc = (row >> 1) + 1, c = row & 1;
            T acc = 0;
            for (int j = 0; j < b; ++j) {
                const cx<T> v = tile[j * P + sc];
                acc += (c ? v.im : v.re) * initch[j * terms + k];
            }
            dst[o] = acc;
        }
    }
    if (threadIdx.x < P) sums_part[threadIdx.x] = c64{sre, sim};
}

// sumsec[sc] = (i/M) * sum over blocks of the partial sums (the reference's
// MPI_Reduce + MPI_Bcast + scaling, src/fmmfft.c:452-457).  Deterministic:
// fixed strided order and a fixed tree.
__global__ __launch_bounds__(256) void k_sums_reduce(Geom g, const c64 *__restrict__ sums_part, int nblk,
                                                     int64_t sums_stride, c64 *__restrict__ sumsec,
                                                     int64_t sumsec_stride) {
    __shared__ double rre[256], rim[256];
    const int sc = blockIdx.x, sig = blockIdx.z;
    sums_part += (int64_t)sig * sums_stride;
    double re = 0.0, im = 0.0;
    for (int i = threadIdx.x; i < nblk; i += blockDim.x) {
        c64 v = sums_part[(int64_t)i * g.P + sc];
        re += v.re;
        im += v.im;
    }
    rre[threadIdx.x] = re;
    rim[threadIdx.x] = im;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            rre[threadIdx.x] += rre[threadIdx.x + h];
            rim[threadIdx.x] += rim[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double inv = 1.0 / (double)g.M;
        sumsec[(int64_t)sig * sumsec_stride + sc] =
            sc == 0 ? c64{0.0, 0.0} : c64{-rim[0] * inv, rre[0] * inv};
    }
}

// ---------------------------------------------------------------------------
// Upward pass, parent level l: phi_l[q] = shftfn . phi_{l+1}[2q] +
// shftfp . phi_{l+1}[2q+1] (src/fmmfft.c:476-515, mpp :174-184).
// One lane per output (node, section, re/im, term).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_m2m(Geom g, int l, T *__restrict__ phi, int64_t tree_stride,
                                             const T *__restrict__ ops) {
    const int terms = g.terms, rows = g.S * 2, node_len = rows * terms;
    const int64_t nout = (int64_t(1) << l) * node_len;
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= nout) return;
    phi += (int64_t)blockIdx.z * tree_stride;
    const int64_t q = o / node_len;
    const int r = (int)(o - q * node_len);
    const int row = r / terms, k = r - row * terms;
    const T *A = ops + ((size_t)(l - 2) * NOPS + OP_SHFTFN) * terms * terms + (size_t)k * terms;
    const T *B = ops + ((size_t)(l - 2) * NOPS + OP_SHFTFP) * terms * terms + (size_t)k * terms;
    const T *cl = phi + level_off(g, l + 1) + (2 * q) * node_len + (int64_t)row * terms;
    const T *cr = cl + node_len;
    T acc = 0;
    for (int kk = 0; kk < terms; ++kk) acc += A[kk] * cl[kk] + B[kk] * cr[kk];
    phi[level_off(g, l) + o] = acc;
}

// ---------------------------------------------------------------------------
// Downward pass at level l.  l == 2: psi_2[q] = topflp . phi_2[q+2 mod 4]
// (src/fmmfft.c:620-639).  l >= 3: psi_l[c] = L2L(psi_{l-1}[c/2]) plus three
// M2L flips from the interaction list (src/fmmfft.c:644-749):
//   even c: shftlp, flip2p(c-2), flip2n(c+2), flip3n(c+3)
//   odd  c: shftln, flip3p(c-3), flip2p(c-2), flip2n(c+2)
// with periodic node indices (the reference's ring of ranks).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_down(Geom g, int l, const T *__restrict__ phi, T *__restrict__ psi,
                                              int64_t tree_stride, const T *__restrict__ ops,
                                              const T *__restrict__ topflp) {
    const int terms = g.terms, rows = g.S * 2, node_len = rows * terms;
    const int nn = 1 << l;
    const int64_t nout = (int64_t)nn * node_len;
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= nout) return;
    phi += (int64_t)blockIdx.z * tree_stride;
    psi += (int64_t)blockIdx.z * tree_stride;
    const int c = (int)(o / node_len);
    const int r = (int)(o - (int64_t)c * node_len);
    const int row = r / terms, k = r - row * terms;
    const T *ph = phi + level_off(g, l) + (int64_t)row * terms;
    const int tt = terms * terms;
    T acc = 0;
    if (l == 2) {
        const T *A = topflp + (size_t)k * terms;
        const T *src = ph + (int64_t)((c + 2) & 3) * node_len;
        for (int kk = 0; kk < terms; ++kk) acc += A[kk] * src[kk];
    } else {
        const T *o8 = ops + (size_t)(l - 3) * NOPS * tt + (size_t)k * terms;
        const T *par = psi + level_off(g, l - 1) + (int64_t)(c >> 1) * node_len + (int64_t)row * terms;
        const T *Ls, *F0, *F1, *F2;
        const T *s0, *s1, *s2;
        if ((c & 1) == 0) {
            Ls = o8 + OP_SHFTLP * tt;
            F0 = o8 + OP_FLIP2P * tt; s0 = ph + (int64_t)pmod(c - 2, nn) * node_len;
            F1 = o8 + OP_FLIP2N * tt; s1 = ph + (int64_t)pmod(c + 2, nn) * node_len;
            F2 = o8 + OP_FLIP3N * tt; s2 = ph + (int64_t)pmod(c + 3, nn) * node_len;
        } else {
            Ls = o8 + OP_SHFTLN * tt;
            F0 = o8 + OP_FLIP3P * tt; s0 = ph + (int64_t)pmod(c - 3, nn) * node_len;
            F1 = o8 + OP_FLIP2P * tt; s1 = ph + (int64_t)pmod(c - 2, nn) * node_len;
            F2 = o8 + OP_FLIP2N * tt; s2 = ph + (int64_t)pmod(c + 2, nn) * node_len;
        }
        for (int kk = 0; kk < terms; ++kk)
            acc += Ls[kk] * par[kk] + F0[kk] * s0[kk] + F1[kk] * s1[kk] + F2[kk] * s2[kk];
    }
    psi[level_off(g, l) + o] = acc;
}

// ---------------------------------------------------------------------------
// Leaf evaluation: psec[mu][sc] = L2T(local expansion of the home box) +
// sum over the 3b near sources of K_sc(nu - mu) x_sc[nu],
// K_sc(d) = -cot(pi (d + sc/P) / M) / M  (src/fmmfft.c:751-954).
// Target mu = i*b + j of section sc sits at mu - sc/P; for j = 0 and
// sc > P/2 its home is box i-1 (row evalmh, sources in boxes i-2..i), for
// sc = P/2 both evaluations are averaged (src/fmmfft.c:771-803, 839-954).
// Column 0 receives x_0 for the combine.
//
// Workgroup: `bb` consecutive boxes x `G` consecutive sections.  The x window
// (boxes i0-2 .. i0+bb) of those sections, their kernel generators and the
// local expansions of boxes i0-1 .. i0+bb-1 are staged in LDS.  A lane owns
// four consecutive targets of one (section, box) and slides the Toeplitz
// kernel through registers: one x load and one kernel load per 4 complex MACs.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_leaf(Geom g, const cx<T> *__restrict__ x, int64_t x_stride,
                                              const T *__restrict__ psi, int64_t tree_stride,
                                              cx<T> *__restrict__ psec, int64_t psec_stride,
                                              const T *__restrict__ evalm, const T *__restrict__ evalmh,
                                              const T *__restrict__ kv, int G, int bb, int conj) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int P = g.P, b = g.b, terms = g.terms, S = g.S;
    const int64_t M = g.M;
    const int rows = (bb + 3) * b;
    const int kvl = 4 * b + 1;  // padded row of the kernel generator
    cx<T> *xw = reinterpret_cast<cx<T> *>(smem);                 // [rows][G]
    T *kvs = reinterpret_cast<T *>(xw + (size_t)rows * G);        // [G][kvl]
    T *loc = kvs + (size_t)G * kvl;                               // [bb+1][G][2][terms]
    const int sig = blockIdx.z;
    x += (int64_t)sig * x_stride;
    psi += (int64_t)sig * tree_stride + level_off(g, g.n);
    psec += (int64_t)sig * psec_stride;
    const int i0 = blockIdx.x * bb;
    const int sc0 = blockIdx.y * G;

    // Stage the x window, kernels and local expansions.
    const int64_t mu_w0 = (int64_t)(i0 - 2) * b;
    for (int e = threadIdx.x; e < rows * G; e += blockDim.x) {
        const int r = e / G, gg = e - r * G;
        int64_t mu = mu_w0 + r;
        mu = mu < 0 ? mu + M : (mu >= M ? mu - M : mu);
        cx<T> v = x[mu * P + sc0 + gg];
        if (conj) v.im = -v.im;
        xw[e] = v;
    }
    for (int e = threadIdx.x; e < G * 4 * b; e += blockDim.x) {
        const int gg = e / (4 * b), d = e - gg * 4 * b;
        const int sc = sc0 + gg;
        kvs[gg * kvl + d] = sc == 0 ? T(0) : kv[(size_t)(sc - 1) * 4 * b + d];
    }
    const int nl = 2 * terms;
    for (int e = threadIdx.x; e < (bb + 1) * G * nl; e += blockDim.x) {
        const int bx = e / (G * nl), rem = e - bx * G * nl;
        const int gg = rem / nl, t = rem - gg * nl;
        const int sc = sc0 + gg;
        const int box = pmod(i0 - 1 + bx, g.s);
        loc[e] = sc == 0 ? T(0) : psi[((int64_t)box * S + (sc - 1)) * nl + t];
    }
    __syncthreads();

    const int nq = b / 4;
    const int items = G * bb * nq;
    for (int w = threadIdx.x; w < items; w += blockDim.x) {
        const int gg = w % G;
        const int rest = w / G;
        const int jq = rest % nq, bl = rest / nq;
        const int sc = sc0 + gg;
        const int j0 = 4 * jq;
        const int64_t mu0 = (int64_t)(i0 + bl) * b + j0;
        if (sc == 0) {
            for (int r = 0; r < 4; ++r) psec[(mu0 + r) * P] = xw[((bl + 2) * b + j0 + r) * G + gg];
            continue;
        }
        const T *kr = kvs + gg * kvl;
        // near field: window rows (bl+1)*b .. (bl+4)*b-1, K index q + b - j0 - r
        T are[4] = {0, 0, 0, 0}, aim[4] = {0, 0, 0, 0};
        const int cidx = b - j0;
        T k0 = kr[cidx], k1 = kr[cidx - 1], k2 = kr[cidx - 2], k3 = kr[cidx - 3];
        const cx<T> *xr = xw + (size_t)(bl + 1) * b * G + gg;
        for (int q = 0; q < 3 * b; ++q) {
            const cx<T> v = xr[(size_t)q * G];
            are[0] += k0 * v.re; aim[0] += k0 * v.im;
            are[1] += k1 * v.re; aim[1] += k1 * v.im;
            are[2] += k2 * v.re; aim[2] += k2 * v.im;
            are[3] += k3 * v.re; aim[3] += k3 * v.im;
            k3 = k2; k2 = k1; k1 = k0;
            k0 = kr[q + 1 + cidx];
        }
        // far field: L2T rows of the home box (loc index bl+1)
        const T *lre = loc + ((size_t)(bl + 1) * G + gg) * nl;
        const T *lim = lre + terms;
        const T *ev = evalm + ((size_t)(sc - 1) * b + j0) * terms;
        for (int r = 0; r < 4; ++r) {
            T fre = 0, fim = 0;
            for (int k = 0; k < terms; ++k) {
                const T e = ev[r * terms + k];
                fre += e * lre[k];
                fim += e * lim[k];
            }
            are[r] += fre;
            aim[r] += fim;
        }
        if (j0 == 0 && 2 * sc >= P) {
            // target 0 as point b of box i-1: sources boxes i-2..i, K index q
            T hre = 0, him = 0;
            const cx<T> *xh = xw + (size_t)bl * b * G + gg;
            for (int q = 0; q < 3 * b; ++q) {
                const cx<T> v = xh[(size_t)q * G];
                hre += kr[q] * v.re;
                him += kr[q] * v.im;
            }
            const T *hl = loc + ((size_t)bl * G + gg) * nl;
            const T *eh = evalmh + (size_t)(sc - P / 2) * terms;
            for (int k = 0; k < terms; ++k) {
                hre += eh[k] * hl[k];
                him += eh[k] * hl[k + terms];
            }
            if (2 * sc == P) {
                are[0] = T(0.5) * (are[0] + hre);
                aim[0] = T(0.5) * (aim[0] + him);
            } else {
                are[0] = hre;
                aim[0] = him;
            }
        }
        for (int r = 0; r < 4; ++r) psec[(mu0 + r) * P + sc] = cx<T>{are[r], aim[r]};
    }
}

// ---------------------------------------------------------------------------
// Combine (src/fmmfft.c:956-970): for every row mu,
//   y_k2[mu] = x_0[mu] + sum_sc w_P^{sc k2} gfac_sc (psec[mu][sc] - sumsec_sc),
// i.e. the reference's dense fo product done as a P-point FFT, in FP64, and
// written block-major (the reference's MPI_Alltoall).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_combine(Geom g, const cx<T> *__restrict__ psec, int64_t psec_stride,
                                                 const c64 *__restrict__ sumsec, int64_t sumsec_stride,
                                                 const c64 *__restrict__ gfac, const c64 *__restrict__ wP,
                                                 cx<T> *__restrict__ y, int64_t y_stride, int rb) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int P = g.P, lp = __builtin_ctz(P);
    const int rs = P + 1;  // padded row
    c64 *a = reinterpret_cast<c64 *>(smem);  // [rb][P+1]
    const int sig = blockIdx.z;
    psec += (int64_t)sig * psec_stride;
    sumsec += (int64_t)sig * sumsec_stride;
    y += (int64_t)sig * y_stride;
    const int64_t mu0 = (int64_t)blockIdx.x * rb;

    for (int e = threadIdx.x; e < rb * P; e += blockDim.x) {
        const int r = e / P, sc = e - r * P;
        const cx<T> v = psec[(mu0 + r) * P + sc];
        c64 w{(double)v.re, (double)v.im};
        if (sc != 0) {
            const c64 s = sumsec[sc];
            w = cmul(gfac[sc], c64{w.re - s.re, w.im - s.im});
        }
        const int br = (int)(__builtin_bitreverse32((unsigned)sc) >> (32 - lp));
        a[r * rs + (lp ? br : 0)] = w;
    }
    __syncthreads();
    for (int h = 1; h < P; h <<= 1) {
        for (int bf = threadIdx.x; bf < rb * (P / 2); bf += blockDim.x) {
            const int r = bf / (P / 2), t = bf - r * (P / 2);
            const int grp = t / h, k = t - grp * h;
            const int i1 = grp * 2 * h + k, i2 = i1 + h;
            const c64 tw = wP[k * (P / (2 * h))];
            const c64 u = a[r * rs + i1];
            const c64 v = cmul(a[r * rs + i2], tw);
            a[r * rs + i1] = c64{u.re + v.re, u.im + v.im};
            a[r * rs + i2] = c64{u.re - v.re, u.im - v.im};
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < rb * P; e += blockDim.x) {
        const int k2 = e / rb, r = e - k2 * rb;
        const c64 v = a[r * rs + k2];
        y[(int64_t)k2 * g.M + mu0 + r] = cx<T>{(T)v.re, (T)v.im};
    }
}

}  // namespace fmmfft
