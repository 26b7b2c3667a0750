// FMM passes of the FMM-FFT on gfx950: S2M with section sums, the upward
// (M2M) and downward (top flip, M2L, L2L) tree passes, the leaf evaluation
// (L2T + near field + half-shift rule) and the section combine.
//
// Layouts (one signal; batched launches add blockIdx.z * stride):
//   x     [M][P]  complex, natural order: section sc is column sc
//                 (the reference's polyphase permute, src/fmmfft.c:419-439,
//                 is only an index map here)
//   tree  levels 2..n, level l has 2^l nodes of [terms][P][2] reals:
//                 term-major, rows (section, re/im) innermost, so lanes that
//                 own consecutive rows access consecutive addresses; the
//                 rows of section 0 are carried (never read) so that a
//                 term row is 2P reals and row blocks stay aligned
//   psec  [P][M]  complex: row sc >= 1 is the cot sum of section sc,
//                 row 0 is a copy of x_0
//   y     [P][M]  complex: block k2 feeds the k2-th length-M FFT
#pragma once
#include "common.hpp"
#include "kernels_fft.hpp"

#include <type_traits>

namespace fmmfft {

// ---------------------------------------------------------------------------
// S2M + per-workgroup section partial sums.
// The reference's initmm (src/fmmfft.c:199-211, called at :465):
//   phi_leaf[box][sc][k] = sum_j x_sc[box*b + j] * initch[j][k]
// and its per-rank section sums (src/fmmfft.c:441-450), accumulated in FP64.
// A tile is BX consecutive boxes of all P sections, i.e. one contiguous
// 64 KB range of x (BX = 64 KB / (b P esize), at least 1 by the plan's
// leaf * P limit).  A persistent workgroup streams its tiles with 16-byte
// loads into registers, one tile ahead, commits a tile to LDS and then
// computes work items (box, section, group of 4 terms): b samples read from
// LDS (lanes = consecutive sections: no bank conflicts) against initch rows
// (broadcast), 8 accumulators per item.  Items with term group 0 also add
// the samples into the FP64 section sums.
// ---------------------------------------------------------------------------
constexpr int kS2mThreads = 512;

template <typename T, int PT, bool TAB = false>
__global__ __launch_bounds__(kS2mThreads) void k_s2m(Geom g, SigAddr<const cx<T>> xa,
                                                     T *__restrict__ phi, int64_t tree_stride,
                                                     c64 *__restrict__ sums_part, int64_t sums_stride,
                                                     const T *__restrict__ initch, int conj) {
    using V = typename std::conditional<sizeof(T) == 4, float4, double2>::type;  // 16 bytes
    constexpr int NT = kS2mThreads;
    constexpr int TILE = 65536 / (int)sizeof(cx<T>);   // complex elements per tile (64 KB)
    constexpr int NV = 65536 / 16, VPER = NV / NT;      // 16-byte vectors
    constexpr int KG = PT / 4;                          // term groups of 4
    __shared__ __align__(16) cx<T> xs[TILE];
    __shared__ double red[2][NT];
    extern __shared__ __align__(16) unsigned char smem[];
    T *ich = reinterpret_cast<T *>(smem);  // [b][PT], zero padded
    const int P = g.P, b = g.b, terms = g.terms;
    const int rows = tree_rows(g);
    const int bx_n = min(TILE / (b * P), g.s);          // boxes per tile
    const int tile_elems = bx_n * b * P, tile_vecs = tile_elems * (int)sizeof(cx<T>) / 16;
    const int ntiles = g.s / bx_n;
    const int sig = blockIdx.z;
    const cx<T> *__restrict__ x = xa.template at<TAB>(sig);
    phi += (int64_t)sig * tree_stride + level_off(g, g.n);
    sums_part += (int64_t)sig * sums_stride + (int64_t)blockIdx.x * P;
    for (int i = threadIdx.x; i < b * PT; i += NT) {
        const int j = i / PT, k = i - j * PT;
        ich[i] = k < terms ? initch[j * terms + k] : T(0);
    }

    V v[VPER];
    auto prefetch = [&](int t) __attribute__((always_inline)) {
        const V *src = reinterpret_cast<const V *>(x + (int64_t)t * tile_elems);
#pragma unroll
        for (int u = 0; u < VPER; ++u) {
            const int e = threadIdx.x + u * NT;
            v[u] = src[min(e, tile_vecs - 1)];
        }
    };
    double sre = 0.0, sim = 0.0;
    const int n_items = bx_n * P * KG;
    int t = blockIdx.x;
    if (t < ntiles) prefetch(t);
    while (t < ntiles) {
        __syncthreads();  // the previous tile is no longer read (and ich is staged)
        V *xv = reinterpret_cast<V *>(xs);
#pragma unroll
        for (int u = 0; u < VPER; ++u) {
            const int e = threadIdx.x + u * NT;
            V w = v[u];
            if (conj) {
                if constexpr (sizeof(T) == 4) {
                    w.y = -w.y;
                    w.w = -w.w;
                } else {
                    w.y = -w.y;
                }
            }
            if (e < tile_vecs) xv[e] = w;
        }
        __syncthreads();
        const int tn = t + gridDim.x;
        if (tn < ntiles) prefetch(tn);
        for (int it = threadIdx.x; it < n_items; it += NT) {
            const int sc = it % P, kg = (it / P) % KG, bx = it / (P * KG);
            const cx<T> *col = xs + bx * b * P + sc;
            const T *ir0 = ich + 4 * kg;
            T are[4] = {0, 0, 0, 0}, aim[4] = {0, 0, 0, 0};
            double pre = 0.0, pim = 0.0;
#pragma unroll 8
            for (int j = 0; j < b; ++j) {
                const cx<T> xv1 = col[j * P];
                const T *ir = ir0 + j * PT;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    are[k] += xv1.re * ir[k];
                    aim[k] += xv1.im * ir[k];
                }
                pre += (double)xv1.re;
                pim += (double)xv1.im;
            }
            if (kg == 0) {
                sre += pre;
                sim += pim;
            }
            {
                T *dst = phi + (int64_t)(t * bx_n + bx) * terms * rows + 2 * sc;
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (4 * kg + k < terms) {
                        dst[(int64_t)(4 * kg + k) * rows] = are[k];
                        dst[(int64_t)(4 * kg + k) * rows + 1] = aim[k];
                    }
            }
        }
        t = tn;
    }
    // per-section partial sums of this workgroup: threads with equal it % P
    // (and term group 0) hold them; fixed-order reduction
    red[0][threadIdx.x] = sre;
    red[1][threadIdx.x] = sim;
    __syncthreads();
    if ((int)threadIdx.x < P) {
        double re = 0.0, im = 0.0;
        for (int q = threadIdx.x; q < NT; q += P) {
            re += red[0][q];
            im += red[1][q];
        }
        sums_part[threadIdx.x] = c64{re, im};
    }
}

// sumsec[sc] = (i/M) * sum over blocks of the partial sums (the reference's
// MPI_Reduce + MPI_Bcast + scaling, src/fmmfft.c:452-457).  Deterministic:
// fixed strided order and a fixed tree.
struct SumsJob {
    const c64 *part;  // [signal][nblk][P] per-workgroup partial sums
    int nblk;
    int64_t part_stride;
    c64 *sec;  // [signal][P + 1]: (i/M) S_sc (0 for sc = 0), then S_0
    int64_t sec_stride;
};
// One 256-thread workgroup's share (section sc of signal sig): a workgroup
// of k_sums_reduce, or one of the first workgroups of the first paired
// upward launch (which saves the reduction its own launch).
__device__ inline void sums_reduce_block(const Geom &g, const SumsJob &j, int sc, int sig) {
    __shared__ double rre[256], rim[256];
    const c64 *__restrict__ sums_part = j.part + (int64_t)sig * j.part_stride;
    const int nblk = j.nblk;
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
        c64 *sec = j.sec + (int64_t)sig * j.sec_stride;
        sec[sc] = sc == 0 ? c64{0.0, 0.0} : c64{-rim[0] * inv, rre[0] * inv};
        if (sc == 0) sec[g.P] = c64{rre[0], rim[0]};  // S_0 itself, for the exact DC bins
    }
}
// (internal linkage: both translation units of the library launch it)
static __global__ __launch_bounds__(256) void k_sums_reduce(Geom g, SumsJob j) {
    sums_reduce_block(g, j, blockIdx.x, blockIdx.z);
}

// ---------------------------------------------------------------------------
// Leaf evaluation: psec[sc][mu] = L2T(local expansion of the home box) +
// sum over the 3b near sources of K_sc(nu - mu) x_sc[nu],
// K_sc(d) = -cot(pi (d + sc/P) / M) / M  (src/fmmfft.c:751-954).
// Target mu = i*b + j of section sc sits at mu - sc/P; for j = 0 and
// sc > P/2 its home is box i-1 (row evalmh, sources in boxes i-2..i), for
// sc = P/2 both evaluations are averaged (src/fmmfft.c:771-803, 839-954).
// Row 0 receives x_0 for the combine.
//
// Vector-FMA form, used when the leaf size is not a multiple of 16.
// Workgroup: `bb` consecutive boxes x `G` consecutive sections.  The x window
// (boxes i0-2 .. i0+bb) of those sections, their kernel generators and the
// local expansions of boxes i0-1 .. i0+bb-1 are staged in LDS.  A lane owns
// four consecutive targets of one (section, box) and slides the Toeplitz
// kernel through registers: one x load and one kernel load per 4 complex MACs.
// ---------------------------------------------------------------------------
template <typename T, bool TAB = false>
__global__ __launch_bounds__(256) void k_leaf(Geom g, SigAddr<const cx<T>> xa,
                                              const T *__restrict__ psi, int64_t tree_stride,
                                              cx<T> *__restrict__ psec, int64_t psec_stride,
                                              const T *__restrict__ evalm, const T *__restrict__ evalmh,
                                              const T *__restrict__ kv, int G, int bb, int conj, RowMap rm) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int P = g.P, b = g.b, terms = g.terms;
    const int64_t M = g.M;
    const int rows = (bb + 3) * b;
    const int kvl = 4 * b + 1;  // padded row of the kernel generator
    cx<T> *xw = reinterpret_cast<cx<T> *>(smem);                 // [rows][G]
    T *kvs = reinterpret_cast<T *>(xw + (size_t)rows * G);        // [G][kvl]
    T *loc = kvs + (size_t)G * kvl;                               // [bb+1][G][2][terms]
    const int sig = blockIdx.z;
    const cx<T> *__restrict__ x = xa.template at<TAB>(sig);
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
        const int cc = t / terms, k = t - cc * terms;
        loc[e] = sc == 0 ? T(0) : psi[((int64_t)box * terms + k) * tree_rows(g) + 2 * sc + cc];
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
            // four consecutive rows stay consecutive (mu0 % 4 = 0, TW >= 4)
            for (int r = 0; r < 4; ++r) psec[rm((int)mu0) + r] = xw[((bl + 2) * b + j0 + r) * G + gg];
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
        for (int r = 0; r < 4; ++r) psec[(int64_t)sc * M + rm((int)mu0) + r] = cx<T>{are[r], aim[r]};
    }
}

// ---------------------------------------------------------------------------
// Leaf evaluation on the matrix cores (leaf size b a multiple of 16).
//
// Same sums as k_leaf.  For one section the near field of 8 consecutive
// boxes is the product
//   D[j][(box, c)] = sum_{q < 3b} K_sc(q - b - j) X[q][(box, c)],
//   X[q][(box, c)] = component c of x_sc[(i0 + box - 1) b + q]
// of the b x 3b Toeplitz block (generated on the fly from the 4b kernel
// samples in LDS) with the 3b x 16 window matrix (8 boxes x re/im), i.e.
// b/16 row tiles of 16x16 MFMA outputs, K = 3b in steps of 4.  The L2T term
// appends K = terms steps with the evaluation rows as A and the boxes' local
// expansions as B.  Target 0 of a section with sc >= P/2 is re-homed to box
// i-1 by adding the difference of the two evaluations (two b-long kernel dot
// products and two L2T rows), in full for sc > P/2 and halved for sc = P/2.
//
// Tile: 8 CT boxes x G sections (one section per wave at a time; CT column
// tiles of 8 boxes share every A value, CT = 2 or 4 for leaf 32 or 16 so
// that a tile carries as many MFMAs as one of leaf 64).  The kernel
// is persistent: a workgroup loops over tiles and issues all global loads of
// its next tile (x window of boxes i0-2 .. i0+BB, kernel samples, local
// expansions of boxes i0-1 .. i0+BB-1) into registers before computing the
// current one, so HBM latency hides behind the MFMAs.  Tiles are ordered so
// that the workgroups of one XCD share x lines.
// MFMA lane layout (16x16x4): lane l supplies A[l%16][l/16] and B[l/16][l%16]
// and receives D[row][l%16] for its four accumulators r = 0..3, with
// row = 4*(l/16) + r for FP32 and l/16 + 4*r for FP64.
// ---------------------------------------------------------------------------
// p ? a : b as a bit select (keeps the compiler from folding selects between
// vector elements into a dynamically indexed extract)
__device__ inline float bsel(int p, float a, float b) {
    const unsigned m = 0u - (unsigned)(p != 0);
    return __uint_as_float((__float_as_uint(a) & m) | (__float_as_uint(b) & ~m));
}
__device__ inline double bsel(int p, double a, double b) {
    const unsigned long long m = 0ull - (unsigned long long)(p != 0);
    return __longlong_as_double((long long)(((unsigned long long)__double_as_longlong(a) & m) |
                                            ((unsigned long long)__double_as_longlong(b) & ~m)));
}

// value of lane l ^ 1 (a DPP quad permutation [1,0,3,2]; no LDS round trip)
__device__ inline float swap_pair(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, true));
}
// value of the lane CTRL rotates in (a DPP row rotation within 16 lanes)
template <int CTRL>
__device__ inline float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
// op (commutative) over the lanes of a wave of equal lane % G (G = 2, 4, 8
// or 16), every lane receiving its class's result: DPP row rotations inside
// each 16-lane row, then the row-pair and half-wave swaps of gfx950 (VALU
// only, no LDS round trip; an LDS-permute chain took one dependent LDS
// latency per step)
template <int G, typename Op>
__device__ inline float lanes_reduce_mod(float v, Op op) {
    static_assert(G == 2 || G == 4 || G == 8 || G == 16, "lane classes of 2, 4, 8 or 16");
    if constexpr (G == 2) v = op(v, dpp_f<0x122>(v));  // row_ror:2
    if constexpr (G <= 4) v = op(v, dpp_f<0x124>(v));  // row_ror:4
    if constexpr (G <= 8) v = op(v, dpp_f<0x128>(v));  // row_ror:8
    const auto r16 = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    v = op(__uint_as_float(r16[0]), __uint_as_float(r16[1]));
    const auto r32 = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return op(__uint_as_float(r32[0]), __uint_as_float(r32[1]));
}
__device__ inline double swap_pair(double v) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)u, 0xB1, 0xF, 0xF, true);
    const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), 0xB1, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <typename T>
struct Mfma;
// Output row of accumulator element r held by lanes with l/16 = kr (both
// layouts measured on gfx950).
template <>
struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ inline v4 mma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int kr, int r) { return 4 * kr + r; }
};
template <>
struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ inline v4 mma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int kr, int r) { return kr + 4 * r; }
};

// ---------------------------------------------------------------------------
// S2M on the matrix cores, complex64 (leaf size 16, 32 or 64; 2P a
// multiple of 16 VEC).
// Same sums as k_s2m: for one box, phi[k][n] = sum_j initch[j][k] B[j][n]
// with B[j][n] real n of row box*b + j of x (row = P complex = 2P reals,
// n = 2 sc + re/im), i.e. D = A B with A[k][j] = initch[j][k] zero-padded
// to 16 MT rows and K = b in KS = b/4 steps.  B comes straight from HBM:
// in K step s lane l loads VEC consecutive reals of row 4s + l/16, and
// element e of that vector is column l%16 of N tile e, so N tile e, column
// c holds real chunk*16VEC + c*VEC + e.  A wave owns one chunk of 16 VEC
// reals of the row for all its boxes (A fragments in registers for the
// whole kernel) and adds its reals into FP64 section sums; no LDS until the
// final per-workgroup reduction of the sums.  Boxes come in sibling pairs
// and the first upward level (M2M to level n-1, src/fmmfft.c:476-515) runs
// on the pair's accumulators before they leave registers; with QUAD (tree
// depth >= 4) a wave takes sibling quads and also forms level n-2 from the
// two parents.
// ---------------------------------------------------------------------------
template <typename T, int VEC>
struct VecOf;
template <> struct VecOf<float, 4> { typedef float4 type; };
template <> struct VecOf<float, 2> { typedef float2 type; };
__device__ inline float velem(const float4 &v, int e) { return e == 0 ? v.x : e == 1 ? v.y : e == 2 ? v.z : v.w; }
__device__ inline float velem(const float2 &v, int e) { return e == 0 ? v.x : v.y; }

template <typename T, int VEC, int MT, int KS, bool QUAD, bool TAB = false>
__global__ __launch_bounds__(256) void k_s2m_mma(Geom g, SigAddr<const cx<T>> xa,
                                                 T *__restrict__ phi, int64_t tree_stride,
                                                 c64 *__restrict__ sums_part, int64_t sums_stride,
                                                 const T *__restrict__ initch, const T *__restrict__ opfrag,
                                                 T *__restrict__ sink, int64_t sink_stride, int conj) {
    using V4 = typename Mfma<T>::v4;
    using VT = typename VecOf<T, VEC>::type;
    constexpr int TP = 16 * MT, KT = 4 * MT, FR = TP * TP;  // the tree's padded terms and K steps
    __shared__ double red[4][16 * VEC];
    const int P = g.P, terms = g.terms, rows = tree_rows(g);
    const int nchunk = 2 * P / (16 * VEC), pper = 4 / nchunk;  // chunks of a row, box pairs per workgroup step
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, col = lane & 15, kr = lane >> 4;
    const int chunk = wave % nchunk, pslot = wave / nchunk;
    const int sig = blockIdx.z;
    const T *xr = reinterpret_cast<const T *>(xa.template at<TAB>(sig));
    T *phi_n = phi + (int64_t)sig * tree_stride + level_off(g, g.n);      // leaves
    T *phi_p = phi + (int64_t)sig * tree_stride + level_off(g, g.n - 1);  // their parents
    sums_part += (int64_t)sig * sums_stride + (int64_t)blockIdx.x * P;
    sink += (int64_t)sig * sink_stride;
    const int n0 = chunk * 16 * VEC + col * VEC;  // first real of this lane's columns
    // A fragments: A[16m + col][4s + kr] = initch[4s + kr][16m + col] (zero past terms)
    T a[MT][KS];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int k = 16 * m + col;
            const T v = initch[(4 * s + kr) * terms + min(k, terms - 1)];
            a[m][s] = k < terms ? v : T(0);
        }
    // M2M to level n-1 on the accumulators in place: in K step s lane (kr, col)
    // holds term 16 (s/4) + 4 kr + s%4 of its columns (the D layout), so the
    // operator fragments are read with that K order: am[c][ko][s] =
    // A_c[16 ko + col][16 (s/4) + 4 kr + s%4] from the fragment-ordered table
    // (element A[i][j] at ((j/4) KO + ko) 64 + i%16 + 16 (j%4))
    // (QUAD: the same for level n-2 from the two parents, am2)
    T am[2][MT][KT], am2[QUAD ? 2 : 1][MT][KT];
#pragma unroll
    for (int lv = 0; lv < (QUAD ? 2 : 1); ++lv) {
        const T *o = opfrag + (size_t)(g.n - 3 - lv) * NOPS * FR;
        const T *op[2] = {o + OP_SHFTFN * FR, o + OP_SHFTFP * FR};
#pragma unroll
        for (int c2 = 0; c2 < 2; ++c2)
#pragma unroll
            for (int ko = 0; ko < MT; ++ko)
#pragma unroll
                for (int s = 0; s < KT; ++s) {
                    const int j = 16 * (s / 4) + 4 * kr + (s % 4);
                    const T v = op[c2][((j / 4) * MT + ko) * 64 + col + 16 * (j % 4)];
                    if (lv == 0) am[c2][ko][s] = v;
                    else if constexpr (QUAD) am2[c2][ko][s] = v;
                }
    }
    // the imaginary parts change sign for the conjugated (inverse) input
    T sgn[VEC];
#pragma unroll
    for (int e = 0; e < VEC; ++e) sgn[e] = (conj && ((n0 + e) & 1)) ? T(-1) : T(1);
    double sum[VEC];
#pragma unroll
    for (int e = 0; e < VEC; ++e) sum[e] = 0.0;

    const int64_t rstride = 2 * (int64_t)P;  // reals per row
    // a wave takes sibling boxes (2q, 2q+1) and then their parent q; two
    // half boxes of loads in flight: the next half loads while the current
    // one computes (the last pair prefetches itself again, keeping every
    // load unconditional)
    constexpr int H = KS / 2;
    auto load = [&](VT *v, int64_t bx, int half) __attribute__((always_inline)) {
        const T *src = xr + (bx * KS * 4 + 4 * H * half + kr) * rstride + n0;
        // 16-byte rows with the nontemporal hint: x is streamed once here
        // (its next reader, the leaf, comes after the tree passes); S2M
        // 0.147 -> 0.139 ms at N = 2^26
#pragma unroll
        for (int s = 0; s < H; ++s) {
            if constexpr (VEC == 4) {
                typedef float f4v __attribute__((ext_vector_type(4)));
                const f4v t = __builtin_nontemporal_load(reinterpret_cast<const f4v *>(src + (int64_t)s * 4 * rstride));
                v[s] = VT{t.x, t.y, t.z, t.w};
            } else {
                v[s] = *reinterpret_cast<const VT *>(src + (int64_t)s * 4 * rstride);
            }
        }
    };
    auto mma_half = [&](V4 (&acc)[MT][VEC], const VT *v, int half) __attribute__((always_inline)) {
#pragma unroll
        for (int s = 0; s < H; ++s)
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                const T bv = velem(v[s], e);
                sum[e] += (double)bv;
#pragma unroll
                for (int m = 0; m < MT; ++m) acc[m][e] = Mfma<T>::mma(a[m][H * half + s], bv, acc[m][e]);
            }
    };
    // node [k][n] of a level (tree rows are the reals of an x row); the
    // padding terms k >= terms go to the sink, so every store is
    // unconditional (a store under a lane condition becomes a branch, and
    // the counter wait after a branch covers the loads in flight)
    auto store = [&](T *node, const V4 (&acc)[MT][VEC]) __attribute__((always_inline)) {
        T *dst = node + n0;
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = 16 * m + Mfma<T>::row(kr, r);
                T *d = dst + (int64_t)k * rows;
#pragma unroll
                for (int e = 0; e < VEC; e += 2) {
                    T *de = k < terms ? d + e : sink;
                    *reinterpret_cast<float2 *>(de) =
                        make_float2(sgn[e] * acc[m][e][r], sgn[e + 1] * acc[m][e + 1][r]);
                }
            }
    };
    auto zero = [&](V4 (&acc)[MT][VEC]) __attribute__((always_inline)) {
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int e = 0; e < VEC; ++e) acc[m][e] = V4{0, 0, 0, 0};
    };
    // parent = shftfn . child 0 + shftfp . child 1, on accumulators in the
    // D layout (the operators were read in that K order)
    auto m2m = [&](V4 (&dst)[MT][VEC], const T (&op)[MT][KT], const V4 (&ch)[MT][VEC])
                   __attribute__((always_inline)) {
#pragma unroll
        for (int s = 0; s < KT; ++s)
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                const T bv = ch[s / 4][e][s % 4];
#pragma unroll
                for (int ko = 0; ko < MT; ++ko) dst[ko][e] = Mfma<T>::mma(op[ko][s], bv, dst[ko][e]);
            }
    };
    const int64_t node_len = (int64_t)terms * rows;
    // one sibling pair (2q, 2q+1): leaves, then parent q into up; va holds
    // the first half of box 2q on entry and the first half of box nb on exit
    auto pair = [&](int64_t q, int64_t nb, VT (&va)[H], VT (&vb)[H], V4 (&up)[MT][VEC])
                    __attribute__((always_inline)) {
        V4 acc[2][MT][VEC];
        load(vb, 2 * q, 1);
        zero(acc[0]);
        mma_half(acc[0], va, 0);
        load(va, 2 * q + 1, 0);
        mma_half(acc[0], vb, 1);
        store(phi_n + 2 * q * node_len, acc[0]);
        load(vb, 2 * q + 1, 1);
        zero(acc[1]);
        mma_half(acc[1], va, 0);
        load(va, nb, 0);
        mma_half(acc[1], vb, 1);
        store(phi_n + (2 * q + 1) * node_len, acc[1]);
        zero(up);
        m2m(up, am[0], acc[0]);
        m2m(up, am[1], acc[1]);
        store(phi_p + q * node_len, up);
    };
    VT va[H], vb[H];
    if constexpr (!QUAD) {
        const int64_t npairs = g.s / 2, step = (int64_t)gridDim.x * pper;
        int64_t q = (int64_t)blockIdx.x * pper + pslot;
        if (q < npairs) load(va, 2 * q, 0);
        while (q < npairs) {
            const int64_t nq = q + step;
            V4 up[MT][VEC];
            pair(q, 2 * (nq < npairs ? nq : q), va, vb, up);  // past the end: this pair again (in L2)
            q = nq;
        }
    } else {
        // sibling quads: two pairs, then their parents' parent (level n-2)
        T *phi_g = phi_p + level_off(g, g.n - 2) - level_off(g, g.n - 1);
        const int64_t nquads = g.s / 4, step = (int64_t)gridDim.x * pper;
        int64_t Q = (int64_t)blockIdx.x * pper + pslot;
        if (Q < nquads) load(va, 4 * Q, 0);
        while (Q < nquads) {
            const int64_t nQ = Q + step;
            V4 up[MT][VEC], gp[MT][VEC];
            pair(2 * Q, 4 * Q + 2, va, vb, up);
            zero(gp);
            m2m(gp, am2[0], up);
            pair(2 * Q + 1, 4 * (nQ < nquads ? nQ : Q), va, vb, up);  // past the end: this quad again
            m2m(gp, am2[1], up);
            store(phi_g + Q * node_len, gp);
            Q = nQ;
        }
    }
    // per-section partial sums of this workgroup, fixed order: lane groups
    // (kr), then the waves of each chunk
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
        sum[e] += __shfl_xor(sum[e], 16, 64);
        sum[e] += __shfl_xor(sum[e], 32, 64);
        if (kr == 0) red[wave][col * VEC + e] = sum[e];
    }
    __syncthreads();
    if ((int)threadIdx.x < P) {
        double t2[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int n = 2 * threadIdx.x + h, ch = n / (16 * VEC), w0 = n % (16 * VEC);
            double acc2 = 0.0;
            for (int w = ch; w < 4; w += nchunk) acc2 += red[w][w0];
            t2[h] = acc2;
        }
        sums_part[threadIdx.x] = c64{t2[0], conj ? -t2[1] : t2[1]};
    }
}

// Exact DC bins of the P blocks (complex64 output): X[M k2] = sum_n x[n]
// e^{-2 pi i n k2 / P} = sum_sc S_sc w_P^{sc k2}, the P-point DFT of the
// section sums S_sc = sum_mu x[mu P + sc], which the S2M kernel accumulates
// in FP64.  In the FMM-FFT those P outputs are the sum over a whole block of
// FP32 cot sums, whose small coherent rounding bias adds up M times (at
// M = 2^20 their rms error was 360 times the other bins'), so they are
// rewritten from the FP64 sums after the last FFT pass (plans of 12 terms
// or more, where the FMM's own truncation error at those bins is far below
// the FP32 rounding).  One workgroup per signal: S_0 as the section-sum
// reduction left it (sumsec[P]), S_sc (sc >= 1) from sumsec = (i/M) S_sc;
// the inverse transform conjugates as the last FFT pass does (its sums are
// of the conjugated input).
template <typename T, bool TAB = false>
__global__ __launch_bounds__(64) void k_dc_exact(Geom g, const c64 *__restrict__ sumsec, int64_t sumsec_stride,
                                                 const c64 *__restrict__ wP, SigAddr<cx<T>> out, int conj) {
    __shared__ c64 S[64], W[64];
    const int sig = blockIdx.z, P = g.P, t = threadIdx.x;
    sumsec += (int64_t)sig * sumsec_stride;
    if (t < P) {
        const c64 ss = sumsec[t];
        const double m = (double)g.M;
        W[t] = wP[t];
        S[t] = t == 0 ? sumsec[P] : c64{m * ss.im, -m * ss.re};
    }
    __syncthreads();
    if (t < P) {
        c64 acc{0.0, 0.0};
#pragma unroll 8
        for (int sc = 0; sc < P; ++sc) acc = cadd(acc, cmul(S[sc], W[(sc * t) & (P - 1)]));
        out.template at<TAB>(sig)[(int64_t)t * g.M] = cx<T>{(T)acc.re, (T)(conj ? -acc.im : acc.im)};
    }
}

// Waves per SIMD the register allocation of a leaf instance is held to:
// three for the FP32 leaf-64 instances, which fit 168 registers without
// spilling (left alone they take ~180 and run two workgroups per CU), four
// for FP32 leaf 32 (128 registers), two for FP64 leaf 64.
template <typename T, int NRT, int G, int KT, int CT>
constexpr int leaf_waves() {
    if (sizeof(T) == 4 && G == 4 && NRT == 4 && CT == 1) return 3;
    if (sizeof(T) == 4 && G == 4 && NRT == 2 && CT == 1) return 4;
    if (sizeof(T) == 8 && G == 4 && NRT == 4 && CT == 1 && KT == 4) return 2;
    return 1;
}

template <typename T, int NRT, int G, int KT, int CT, bool TAB = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(leaf_waves<T, NRT, G, KT, CT>())))
void k_leaf_mma(Geom g, SigAddr<const cx<T>> xa,
                                                  const T *__restrict__ psi, int64_t tree_stride,
                                                  cx<T> *__restrict__ psec, int64_t psec_stride,
                                                  const T *__restrict__ evfrag, const T *__restrict__ evalm,
                                                  const T *__restrict__ evalmh, const T *__restrict__ kv, int conj,
                                                  RowMap rm, int nsig) {
    using V4 = typename Mfma<T>::v4;
    constexpr int BB = 8 * CT;       // boxes per tile: CT column tiles of 8 boxes x re/im
    constexpr int b = 16 * NRT;      // leaf size
    constexpr int ROWS = (BB + 3) * b;
    constexpr int XN = ROWS * G, XPER = (XN + 255) / 256;           // x window
    constexpr int KVN = G * 4 * b;                                   // kernel samples
    // local expansions in LDS with the terms padded to TP = 4 KT (the padding
    // slots hold copies of the last term and meet zero L2T rows), so every
    // index below is a compile-time shift or mask
    constexpr int TP = 4 * KT;
    constexpr int NL = TP * 2 * G;                                   // one box
    constexpr int LSTR = NL + 2;                                     // +2: bank shift
    constexpr int LN = (BB + 1) * NL, LPER = (LN + 255) / 256;
    // evaluation rows of the target-0 correction: [G][2][TP] (row of the
    // re-homed target, row of target 0; zero past terms)
    constexpr int EN = G * 2 * TP;
    extern __shared__ __align__(16) unsigned char smem[];
    const int P = g.P, terms = g.terms;
    const int64_t M = g.M;
    const int rowsS = tree_rows(g);      // tree row stride
    // offsets into x and psec fit 32 bits (N <= 2^30; tree offsets are
    // 64-bit); P, M and s are powers of two
    const int lp = __builtin_ctz(P), mmask = (int)M - 1, smask = g.s - 1;
    // window rows of box q start at q b G + boxpad(q): for 8 consecutive
    // boxes the pads cover 8 distinct bank pairs apart from the row stride's
    // bank, so the B loads (lanes over 8 boxes x re/im x 4 rows) are free of
    // LDS bank conflicts
    auto boxpad = [](int q) { return (q & 3) + 8 * (q >> 2); };
    constexpr int XPAD = ((BB + 2) & 3) + 8 * ((BB + 2) >> 2) + 1;  // boxpad(BB + 2) + 1
    cx<T> *xw = reinterpret_cast<cx<T> *>(smem);           // [ROWS][G] + pads
    T *kvs = reinterpret_cast<T *>(xw + XN + XPAD);        // [G][4b]
    T *lsx = kvs + KVN;                                    // [BB+1][TP][G][2] (+2 per box)
    T *evs = lsx + (BB + 1) * LSTR;                        // [G][2][TP]
    psi += level_off(g, g.n);

    // Persistent tiles over all nsig signals (tile = signal * ntiles + lid).
    // Iteration k of workgroup w takes tile k*nwg + xo(w); xo makes the
    // workgroups of one XCD take consecutive tiles (the section groups of one
    // box chunk share x lines).  The grid is a multiple of the number of
    // section groups, so a workgroup keeps the section group xo % ngroups for
    // all its tiles (of every signal) and loads the section constants once.
    const int ngroups = P / G;
    const int ntiles = ngroups * (g.s / BB);
    const int ntot = nsig * ntiles;
    const int nwg = gridDim.x, w = blockIdx.x;
    const int xo = (nwg % 8 == 0) ? (w % 8) * (nwg / 8) + w / 8 : w;
    const int sc0 = (xo % ngroups) * G;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = lane & 15, kr = lane >> 4;
    const int bx = col >> 1, c = col & 1;
    const T *xwc = reinterpret_cast<const T *>(xw);
    static_assert(G <= 4, "one section per wave");
    const int gg = wave, sc = sc0 + gg;

    // Section constants (clamped section index: section 0 and the waves past
    // G = 2 load a valid row they never use): kernel samples [G][4b] and the
    // target-0 correction rows [G][2][TP] (zero past terms) in LDS, and this
    // wave's L2T A fragments (lane order, KT steps per row tile, zero past
    // terms) in registers.  The first tile's barriers publish the LDS part.
    for (int e = threadIdx.x; e < KVN; e += 256) {
        const int sce = min(max(sc0 + e / (4 * b), 1), P - 1);
        kvs[e] = kv[(sce - 1) * 4 * b + e % (4 * b)];
    }
    for (int e = threadIdx.x; e < EN; e += 256) {
        const int sce = min(max(sc0 + e / (2 * TP), 1), P - 1), kk = e % TP;
        const T *row = ((e / TP) & 1) ? evalm + (size_t)(sce - 1) * b * terms
                                      : evalmh + (size_t)max(sce - P / 2, 0) * terms;
        const T v = row[min(kk, terms - 1)];
        evs[e] = kk < terms ? v : T(0);
    }
    T la[NRT][KT];
    {
        const int scl = min(max(sc, 1), P - 1);
        const T *ef = evfrag + (((scl - 1) * NRT * KT) << 6) + lane;
#pragma unroll
        for (int s2 = 0; s2 < KT; ++s2)
#pragma unroll
            for (int rt = 0; rt < NRT; ++rt) la[rt][s2] = ef[(rt * KT + s2) << 6];
    }

    cx<T> xr[XPER];
    T lr[LPER];
    // issue every global load of a tile into registers (they stay in flight
    // while the previous tile computes).  Every load is unconditional
    // (indices clamped): a lane-conditional load becomes a branch, and the
    // counter wait after a branch covers every load in flight.
    auto prefetch = [&](int tile) __attribute__((always_inline)) {
        const int sig = tile / ntiles, lid = tile - sig * ntiles;
        const cx<T> *xs = xa.template at<TAB>(sig);
        const T *ps = psi + (int64_t)sig * tree_stride;
        const int i0 = (lid / ngroups) * BB;
        const int mu_w0 = (i0 - 2) * b;
#pragma unroll
        for (int u = 0; u < XPER; ++u) {
            const int e = (XN % 256 == 0) ? threadIdx.x + 256 * u : min((int)threadIdx.x + 256 * u, XN - 1);
            const int r = e / G, ge = e % G;
            const int mu = (mu_w0 + r) & mmask;  // periodic window
            xr[u] = xs[(mu << lp) + sc0 + ge];
        }
#pragma unroll
        for (int u = 0; u < LPER; ++u) {
            const int e = (LN % 256 == 0) ? threadIdx.x + 256 * u : min((int)threadIdx.x + 256 * u, LN - 1);
            const int bx9 = e / NL, rem = e % NL;
            const int kk = rem / (2 * G), rr = rem % (2 * G);
            const int sce = sc0 + (rr >> 1);
            const int box = (i0 - 1 + bx9) & smask;
            // padding terms meet zero L2T rows; section 0 reads its own rows
            // and never uses them
            lr[u] = ps[((int64_t)box * terms + min(kk, terms - 1)) * rowsS + 2 * sce + (rr & 1)];
        }
    };
    auto commit = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int u = 0; u < XPER; ++u) {
            const int e = threadIdx.x + 256 * u;
            if (XN % 256 == 0 || e < XN) {
                cx<T> v = xr[u];
                if (conj) v.im = -v.im;
                xw[e + boxpad((e / G) / b)] = v;
            }
        }
#pragma unroll
        for (int u = 0; u < LPER; ++u) {
            const int e = threadIdx.x + 256 * u;
            if (LN % 256 == 0 || e < LN) lsx[(e / NL) * LSTR + e % NL] = lr[u];
        }
    };

    int k = 0;
    int tile = xo;
    if (tile < ntot) prefetch(tile);
    while (tile < ntot) {
        __syncthreads();  // the previous tile is no longer read
        commit();
        __syncthreads();
        const int sig = tile / ntiles, lid = tile - sig * ntiles;
        cx<T> *psec_s = psec + (int64_t)sig * psec_stride;
        const int i0 = (lid / ngroups) * BB;
        // past the last tile the prefetch re-reads the current one (still in
        // L2; an unconditional prefetch keeps the counter waits exact)
        const int ntile = (++k) * nwg + xo;
        prefetch(ntile < ntot ? ntile : tile);

        if (gg < G) {
            if (sc == 0) {
                for (int t = lane; t < BB * b; t += 64)
                    psec_s[rm(i0 * b + t)] = xw[(2 * b + t) * G + boxpad((2 * b + t) / b)];
            } else {
                const T *kr_ = kvs + gg * 4 * b;
                V4 acc[NRT][CT];
#pragma unroll
                for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
                    for (int ct = 0; ct < CT; ++ct) acc[rt][ct] = V4{0, 0, 0, 0};
                // near field.  K step s covers window row q = 4 s + kr of box bx
                // (rows (bx+1) b + q); A(rt, s) = K index 4 s + kr + b - 16 rt - col.
                // The Toeplitz structure gives A(rt+1, s+4) = A(rt, s), so each
                // group of four steps loads only the rt = 0 values and shifts the rest.
                constexpr int NG = 3 * b / 16;  // groups of four K steps
                T areg[NRT][4];
#pragma unroll
                for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
                    for (int t = 0; t < 4; ++t) areg[rt][t] = kr_[4 * t + kr + b - 16 * rt - col];
                // window rows of group gi for column tile ct (boxes 8 ct ..):
                // (bx + 8 ct + 1) b + 16 gi + 4 t + kr (+ box pad)
                auto xrow = [&](int gi, int ct) {
                    const int q = bx + 8 * ct + 1;
                    return xwc + (((size_t)q * b + 16 * gi) * G + gg + boxpad(q + gi / (b / 16))) * 2 + c;
                };
                // B loads run one group ahead of the MFMAs that use them; the A
                // values serve all CT column tiles
                T bvc[CT][4];
#pragma unroll
                for (int ct = 0; ct < CT; ++ct) {
                    const T *xg = xrow(0, ct);
#pragma unroll
                    for (int t = 0; t < 4; ++t) bvc[ct][t] = xg[(size_t)(4 * t + kr) * G * 2];
                }
#pragma unroll
                for (int gi = 0; gi < NG; ++gi) {
                    T bvn[CT][4], an[4];
                    if (gi + 1 < NG) {
#pragma unroll
                        for (int ct = 0; ct < CT; ++ct) {
                            const T *xg = xrow(gi + 1, ct);
#pragma unroll
                            for (int t = 0; t < 4; ++t) bvn[ct][t] = xg[(size_t)(4 * t + kr) * G * 2];
                        }
#pragma unroll
                        for (int t = 0; t < 4; ++t) an[t] = kr_[16 * (gi + 1) + 4 * t + kr + b - col];
                    }
                    // keep the next group's loads ahead of this group's MFMAs
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int t = 0; t < 4; ++t)
#pragma unroll
                        for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
                            for (int ct = 0; ct < CT; ++ct)
                                acc[rt][ct] = Mfma<T>::mma(areg[rt][t], bvc[ct][t], acc[rt][ct]);
                    if (gi + 1 < NG) {
#pragma unroll
                        for (int t = 0; t < 4; ++t) {
#pragma unroll
                            for (int rt = NRT - 1; rt > 0; --rt) areg[rt][t] = areg[rt - 1][t];
                            areg[0][t] = an[t];
#pragma unroll
                            for (int ct = 0; ct < CT; ++ct) bvc[ct][t] = bvn[ct][t];
                        }
                    }
                }
                // far field (L2T): B = local expansion of box bx + 8 ct from LDS
#pragma unroll
                for (int s2 = 0; s2 < KT; ++s2)
#pragma unroll
                    for (int ct = 0; ct < CT; ++ct) {
                        const T bvl = lsx[(bx + 8 * ct + 1) * LSTR + 2 * gg + c + (4 * s2 + kr) * 2 * G];
#pragma unroll
                        for (int rt = 0; rt < NRT; ++rt) acc[rt][ct] = Mfma<T>::mma(la[rt][s2], bvl, acc[rt][ct]);
                    }
                if (2 * sc >= P) {
                    // re-home target 0 of box q = bx + 8 ct to box q - 1: lanes
                    // col, col+16, col+32, col+48 each take a quarter of the terms
#pragma unroll
                    for (int ct = 0; ct < CT; ++ct) {
                        const int q = bx + 8 * ct;
                        T corr = 0;
                        const T *xm2 = xwc + ((size_t)q * b * G + gg + boxpad(q)) * 2 + c;              // box i-2
                        const T *xp1 = xwc + ((size_t)(q + 3) * b * G + gg + boxpad(q + 3)) * 2 + c;    // box i+1
#pragma unroll 4
                        for (int j = 0; j < b / 4; ++j) {
                            const int t = 4 * j + kr;
                            corr += kr_[t] * xm2[(size_t)t * G * 2] - kr_[3 * b + t] * xp1[(size_t)t * G * 2];
                        }
                        const T *lm = lsx + q * LSTR + 2 * gg + c;        // box i-1
                        const T *lb = lsx + (q + 1) * LSTR + 2 * gg + c;  // box i
                        const T *eh = evs + gg * 2 * TP, *e0 = eh + TP;
#pragma unroll
                        for (int j = 0; j < KT; ++j)
                            corr += eh[4 * j + kr] * lm[(4 * j + kr) * 2 * G] - e0[4 * j + kr] * lb[(4 * j + kr) * 2 * G];
                        corr += __shfl_xor(corr, 16, 64);
                        corr += __shfl_xor(corr, 32, 64);
                        if (kr == 0) acc[0][ct][0] += (2 * sc == P) ? T(0.5) * corr : corr;
                    }
                }
                // write: pair lanes (c = 0, 1) swap halves so that each lane holds
                // two complete complex outputs; 8 lanes cover 16 consecutive targets
                // (bit selects: a ?: between vector elements becomes an indexed
                // extract over all accumulators, a long compare/select chain)
#pragma unroll
                for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                for (int rt = 0; rt < NRT; ++rt) {
                    cx<T> *dst = psec_s + (int64_t)sc * M;
                    const int mu = (i0 + bx + 8 * ct) * b + 16 * rt;
                    const T a0 = acc[rt][ct][0], a1 = acc[rt][ct][1], a2 = acc[rt][ct][2], a3 = acc[rt][ct][3];
                    const T keep0 = bsel(c, a2, a0), keep1 = bsel(c, a3, a1);
                    const T recv0 = swap_pair(bsel(c, a0, a2));
                    const T recv1 = swap_pair(bsel(c, a1, a3));
                    const cx<T> o0{bsel(c, recv0, keep0), bsel(c, keep0, recv0)};
                    const cx<T> o1{bsel(c, recv1, keep1), bsel(c, keep1, recv1)};
                    if constexpr (sizeof(T) == 4) {
                        // rows 4 kr + 2c, 4 kr + 2c + 1: one 16-byte store (an
                        // even pair stays a pair in the row map, TW >= 4)
                        *reinterpret_cast<float4 *>(dst + rm(mu + Mfma<T>::row(kr, 2 * c))) =
                            make_float4(o0.re, o0.im, o1.re, o1.im);
                    } else {
                        dst[rm(mu + Mfma<T>::row(kr, 2 * c))] = o0;
                        dst[rm(mu + Mfma<T>::row(kr, 2 * c + 1))] = o1;
                    }
                }
            }
        }
        tile = ntile;
    }
}

// ---------------------------------------------------------------------------
// Leaf evaluation for complex64 on the FP16 matrix pipe (leaf 32 or 64).
//
// Same sums and tiles as k_leaf_mma (8 boxes x G sections per tile, one
// section per wave, persistent), with the near field D = K X run as FP16
// split products on v_mfma_f32_16x16x16_f16 with FP32 accumulation:
//   K = Kh + Kl (2^14 K_sc split into two FP16 parts at plan time, kfrag),
//   X = s (Xh + Xl) (the wave's window scaled by a power of two s so that
//       its largest |value| lies in [2^14, 2^15), split on the fly),
//   D = (Kh Xh + Kh Xl + Kl Xh) / (2^14 s),
// each operand carried to ~2^-22 and only the Kl Xl term (~2^-22) dropped.
// The singular near-diagonal samples d = nu - mu in {0, -1} (|d + sc/P| <
// 1, the largest |K| by far: up to P/pi) are not in the FP16 product: they
// are added per target in FP64 from the exact FP32 window and the FP64
// kernel values.  L2T (K = terms) and the target-0 half-shift correction
// stay FP32 as in k_leaf_mma (src/fmmfft.c:751-954).
//
// A fragments (Toeplitz: A(rt, ks) = F(ks - rt), NF = 3b/16 + NRT - 1 of
// them) stay in registers for the whole kernel (the wave's section never
// changes).  The window sits in LDS as FP32, one array per (section,
// component) with a pad of 8 floats per box, so that the B operand of K
// step ks (4 consecutive window rows of one box and component per lane) is
// one conflict-free 16-byte read: box stride 4(b+8) bytes is an odd
// multiple of 32 modulo 256 and the component arrays differ by 16 modulo
// 256.  Loads of the next tile are issued into registers before the current
// tile computes, in units of 4 rows of one section (written to LDS as one
// 16-byte store per component); the per-section maxima for the scales come
// from the same registers.
// ---------------------------------------------------------------------------
typedef _Float16 h16x4 __attribute__((ext_vector_type(4)));

template <int NRT>
struct LeafH16 {
    static constexpr int b = 16 * NRT;
    static constexpr int BB = 8, NBOX = BB + 3, ROWS = NBOX * b;
    static constexpr int BSTR = b + 8;                                   // floats per window box
    static constexpr int WSTR = NBOX * BSTR + ((4 - (NBOX * BSTR) % 64) % 64 + 64) % 64;  // = 4 mod 64
    static constexpr int NKS = 3 * b / 16;                               // K steps of the near field
    static constexpr int NF = NKS + NRT - 1;                             // Toeplitz fragments
};

template <int NRT, int G, int KT, bool TAB = false>
__global__ __launch_bounds__(64 * (G > 4 ? G : 4)) __attribute__((amdgpu_waves_per_eu(2)))
void k_leaf_h16(Geom g, SigAddr<const c32> xa, const float *__restrict__ psi,
                int64_t tree_stride, c32 *__restrict__ psec, int64_t psec_stride,
                const float *__restrict__ evfrag, const float *__restrict__ evalm,
                const float *__restrict__ evalmh, const float *__restrict__ kv,
                const h16x4 *__restrict__ kfrag, const double *__restrict__ kv64, int conj, RowMap rm,
                int nsig) {
    using V4 = Mfma<float>::v4;
    using L = LeafH16<NRT>;
    constexpr int b = L::b, BB = L::BB, ROWS = L::ROWS, BSTR = L::BSTR, WSTR = L::WSTR;
    constexpr int NKS = L::NKS, NF = L::NF;
    constexpr int NU = ROWS / 4 * G;                                   // 4-row units of the window
    constexpr int KVN = G * 4 * b;
    constexpr int TP = 4 * KT;
    constexpr int NL = TP * 2 * G;
    constexpr int LSTR = NL + 2;
    constexpr int NWV = G > 4 ? G : 4, NTH = 64 * NWV;              // one section per wave
    constexpr int LN = (BB + 1) * NL, LPER = (LN + NTH - 1) / NTH;
    constexpr int EN = G * 2 * TP;
    constexpr int UPER = (NU + NTH - 1) / NTH;
    static_assert(NTH % G == 0, "a thread's units share one section");
    extern __shared__ __align__(16) unsigned char smem[];
    const int P = g.P, terms = g.terms;
    const int64_t M = g.M;
    const int rowsS = tree_rows(g);
    const int lp = __builtin_ctz(P), mmask = (int)M - 1, smask = g.s - 1;
    float *xw = reinterpret_cast<float *>(smem);  // [2G][WSTR]: (section, re/im) arrays
    float *kvs = xw + 2 * G * WSTR;                // [G][4b]
    float *lsx = kvs + KVN;                        // [BB+1][LSTR]
    float *evs = lsx + (BB + 1) * LSTR;            // [G][2][TP]
    float *smax = evs + EN;                        // [NWV waves][G]
    psi += level_off(g, g.n);

    const int ngroups = P / G;
    const int ntiles = ngroups * (g.s / BB);
    const int ntot = nsig * ntiles;
    const int nwg = gridDim.x, w = blockIdx.x;
    const int xo = (nwg % 8 == 0) ? (w % 8) * (nwg / 8) + w / 8 : w;
    const int sc0 = (xo % ngroups) * G;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = lane & 15, kr = lane >> 4;
    const int bx = col >> 1, c = col & 1;
    const int gg = wave, sc = sc0 + gg;

    for (int e = threadIdx.x; e < KVN; e += NTH) {
        const int sce = min(max(sc0 + e / (4 * b), 1), P - 1);
        kvs[e] = kv[(sce - 1) * 4 * b + e % (4 * b)];
    }
    for (int e = threadIdx.x; e < EN; e += NTH) {
        const int sce = min(max(sc0 + e / (2 * TP), 1), P - 1), kk = e % TP;
        const float *row = ((e / TP) & 1) ? evalm + (size_t)(sce - 1) * b * terms
                                          : evalmh + (size_t)max(sce - P / 2, 0) * terms;
        const float v = row[min(kk, terms - 1)];
        evs[e] = kk < terms ? v : 0.f;
    }
    // this wave's section constants: L2T fragments (FP32), the FP16 split
    // near-field fragments, and the two singular kernel values in FP64
    const int scl = min(max(sc, 1), P - 1);
    float la[NRT][KT];
    {
        const float *ef = evfrag + (((scl - 1) * NRT * KT) << 6) + lane;
#pragma unroll
        for (int s2 = 0; s2 < KT; ++s2)
#pragma unroll
            for (int rt = 0; rt < NRT; ++rt) la[rt][s2] = ef[(rt * KT + s2) << 6];
    }
    h16x4 ah[NF], al[NF];
    {
        const h16x4 *kf = kfrag + (size_t)(scl - 1) * NF * 2 * 64 + lane;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            ah[f] = kf[(2 * f) * 64];
            al[f] = kf[(2 * f + 1) * 64];
        }
    }
    const double k0 = kv64[(size_t)(scl - 1) * 4 * b + 2 * b];      // K_sc(0)
    const double k1 = kv64[(size_t)(scl - 1) * 4 * b + 2 * b - 1];  // K_sc(-1)

    float xr[UPER][8];  // 4 rows x (re, im) per unit
    float lr[LPER];
    // byte offsets of this lane's loads from a tile's first window row and
    // first expansion box (fixed for the kernel)
    uint32_t bx_off[UPER], bl_off[LPER];
#pragma unroll
    for (int u = 0; u < UPER; ++u) {
        const int e = (NU % NTH == 0) ? threadIdx.x + NTH * u : min((int)threadIdx.x + NTH * u, NU - 1);
        const int rq = e / G, ge = e % G;
        bx_off[u] = (uint32_t)(((4 * rq) << lp) + ge) * (uint32_t)sizeof(c32);
    }
#pragma unroll
    for (int u = 0; u < LPER; ++u) {
        const int e = (LN % NTH == 0) ? threadIdx.x + NTH * u : min((int)threadIdx.x + NTH * u, LN - 1);
        const int bx9 = e / NL, rem = e % NL;
        const int kk = rem / (2 * G), rr = rem % (2 * G);
        const int sce = sc0 + (rr >> 1);
        bl_off[u] = (uint32_t)((bx9 * terms + min(kk, terms - 1)) * rowsS + 2 * sce + (rr & 1)) * (uint32_t)sizeof(float);
    }
    auto prefetch = [&](int tile) __attribute__((always_inline)) {
        const int sig = tile / ntiles, lid = tile - sig * ntiles;
        const c32 *xs = xa.template at<TAB>(sig);
        const float *ps = psi + (int64_t)sig * tree_stride;
        const int i0 = (lid / ngroups) * BB;
        const int mu_w0 = (i0 - 2) * b;
        // tiles whose window and expansions do not wrap around the signal
        // (all but the first and last box chunk): one per-tile base and the
        // fixed lane offsets, so an address is one add (-8 us at N = 2^26);
        // the other two take the periodic index row by row
        if (i0 >= 2 && i0 + BB + 1 <= g.s) {
            const c32 *xb = xs + ((int64_t)mu_w0 << lp) + sc0;
            const char *pb = reinterpret_cast<const char *>(ps + (int64_t)(i0 - 1) * terms * rowsS);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const char *xk = reinterpret_cast<const char *>(xb + ((int64_t)k << lp));
#pragma unroll
                for (int u = 0; u < UPER; ++u) {
                    const c32 v = *reinterpret_cast<const c32 *>(xk + bx_off[u]);
                    xr[u][2 * k] = v.re;
                    xr[u][2 * k + 1] = v.im;
                }
            }
#pragma unroll
            for (int u = 0; u < LPER; ++u) lr[u] = *reinterpret_cast<const float *>(pb + bl_off[u]);
            return;
        }
#pragma unroll
        for (int u = 0; u < UPER; ++u) {
            const int e = (NU % NTH == 0) ? threadIdx.x + NTH * u : min((int)threadIdx.x + NTH * u, NU - 1);
            const int rq = e / G, ge = e % G;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int mu = (mu_w0 + 4 * rq + k) & mmask;  // periodic window
                const c32 v = xs[(mu << lp) + sc0 + ge];
                xr[u][2 * k] = v.re;
                xr[u][2 * k + 1] = v.im;
            }
        }
#pragma unroll
        for (int u = 0; u < LPER; ++u) {
            const int e = (LN % NTH == 0) ? threadIdx.x + NTH * u : min((int)threadIdx.x + NTH * u, LN - 1);
            const int bx9 = e / NL, rem = e % NL;
            const int kk = rem / (2 * G), rr = rem % (2 * G);
            const int sce = sc0 + (rr >> 1);
            const int box = (i0 - 1 + bx9) & smask;
            lr[u] = ps[((int64_t)box * terms + min(kk, terms - 1)) * rowsS + 2 * sce + (rr & 1)];
        }
    };
    // window rows to LDS (imaginary parts negated for the inverse) and the
    // per-section maximum magnitude of this wave's units to smax
    auto commit = [&]() __attribute__((always_inline)) {
        float m = 0.f;
        const float sg = conj ? -1.f : 1.f;
#pragma unroll
        for (int u = 0; u < UPER; ++u) {
            const int e = threadIdx.x + NTH * u;
            if (NU % NTH == 0 || e < NU) {
                const int rq = e / G, ge = e % G;
                const int q = (4 * rq) / b, j = (4 * rq) % b;
                float *dre = xw + (2 * ge) * WSTR + q * BSTR + j;
                *reinterpret_cast<float4 *>(dre) = make_float4(xr[u][0], xr[u][2], xr[u][4], xr[u][6]);
                *reinterpret_cast<float4 *>(dre + WSTR) =
                    make_float4(sg * xr[u][1], sg * xr[u][3], sg * xr[u][5], sg * xr[u][7]);
#pragma unroll
                for (int k = 0; k < 8; ++k) m = fmaxf(m, fabsf(xr[u][k]));
            }
        }
        // lanes of equal section: lane % G
        m = lanes_reduce_mod<G>(m, [](float a, float b2) { return fmaxf(a, b2); });
        if (lane < G) smax[wave * G + lane] = m;
#pragma unroll
        for (int u = 0; u < LPER; ++u) {
            const int e = threadIdx.x + NTH * u;
            if (LN % NTH == 0 || e < LN) lsx[(e / NL) * LSTR + e % NL] = lr[u];
        }
    };

    int k = 0;
    int tile = xo;
    if (tile < ntot) prefetch(tile);
    while (tile < ntot) {
        __syncthreads();  // the previous tile is no longer read
        commit();
        __syncthreads();
        const int sig = tile / ntiles, lid = tile - sig * ntiles;
        c32 *psec_s = psec + (int64_t)sig * psec_stride;
        const int i0 = (lid / ngroups) * BB;
        const int ntile = (++k) * nwg + xo;
        prefetch(ntile < ntot ? ntile : tile);

        if (gg < G) {
            const float *xre = xw + (2 * gg) * WSTR, *xim = xre + WSTR;
            if (sc == 0) {
                for (int t = lane; t < BB * b; t += 64) {
                    const int o = (2 + t / b) * BSTR + t % b;
                    psec_s[rm(i0 * b + t)] = c32{xre[o], xim[o]};
                }
            } else {
                // block scale of this section's window: max |value| -> [2^14, 2^15)
                float mx = smax[gg];
#pragma unroll
                for (int w2 = 1; w2 < NWV; ++w2) mx = fmaxf(mx, smax[w2 * G + gg]);
                int ex = 0;
                (void)frexpf(mx, &ex);
                const int es = mx > 0.f ? min(max(15 - ex, -126), 127) : 0;
                const float sx = ldexpf(1.f, es), us = ldexpf(1.f, -es - 14);
                const float *xc = c ? xim : xre;  // this lane's component
                V4 acc[NRT];
#pragma unroll
                for (int rt = 0; rt < NRT; ++rt) acc[rt] = V4{0, 0, 0, 0};
                // near field: K step ks takes window rows (bx+1) b + 16 ks + 4 kr .. +3
                // (4 rows never cross a box).  The window rows of step ks + 1
                // are read before the MFMAs of step ks, and the MFMAs of the
                // NRT row tiles run interleaved (independent accumulators back
                // to back; one tile's three products in a row wait on each
                // other).
                auto brow = [&](int ks) __attribute__((always_inline)) {
                    const int wr = 16 * ks + 4 * kr;
                    return *reinterpret_cast<const float4 *>(xc + (bx + 1 + wr / b) * BSTR + wr % b);
                };
                float4 vn = brow(0);
#pragma unroll
                for (int ks = 0; ks < NKS; ++ks) {
                    const float4 v = vn;
                    if (ks + 1 < NKS) vn = brow(ks + 1);
                    const float xv[4] = {v.x * sx, v.y * sx, v.z * sx, v.w * sx};
                    h16x4 xh, xl;
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        xh[i] = (_Float16)xv[i];
                        xl[i] = (_Float16)(xv[i] - (float)xh[i]);
                    }
#pragma unroll
                    for (int rt = 0; rt < NRT; ++rt)
                        acc[rt] = __builtin_amdgcn_mfma_f32_16x16x16f16(ah[ks - rt + NRT - 1], xh, acc[rt], 0, 0, 0);
#pragma unroll
                    for (int rt = 0; rt < NRT; ++rt)
                        acc[rt] = __builtin_amdgcn_mfma_f32_16x16x16f16(ah[ks - rt + NRT - 1], xl, acc[rt], 0, 0, 0);
#pragma unroll
                    for (int rt = 0; rt < NRT; ++rt)
                        acc[rt] = __builtin_amdgcn_mfma_f32_16x16x16f16(al[ks - rt + NRT - 1], xh, acc[rt], 0, 0, 0);
                }
#pragma unroll
                for (int rt = 0; rt < NRT; ++rt) acc[rt] *= us;
                // far field (L2T) in FP32
#pragma unroll
                for (int s2 = 0; s2 < KT; ++s2) {
                    const float bvl = lsx[(bx + 1) * LSTR + 2 * gg + c + (4 * s2 + kr) * 2 * G];
#pragma unroll
                    for (int rt = 0; rt < NRT; ++rt) acc[rt] = Mfma<float>::mma(la[rt][s2], bvl, acc[rt]);
                }
                // singular samples in FP64: K(0) x[mu] + K(-1) x[mu-1] for the
                // lane's targets mu = 16 rt + 4 kr + r of box bx + 2
#pragma unroll
                for (int rt = 0; rt < NRT; ++rt) {
                    const int j0 = 16 * rt + 4 * kr;
                    const float4 v = *reinterpret_cast<const float4 *>(xc + (bx + 2) * BSTR + j0);
                    const float vp = xc[j0 == 0 ? (bx + 1) * BSTR + b - 1 : (bx + 2) * BSTR + j0 - 1];
                    const float xs4[5] = {vp, v.x, v.y, v.z, v.w};
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        acc[rt][r] += (float)(k0 * (double)xs4[r + 1] + k1 * (double)xs4[r]);
                }
                if (2 * sc >= P) {
                    // re-home target 0 of box bx to box bx - 1 (as k_leaf_mma)
                    const float *kr_ = kvs + gg * 4 * b;
                    float corr = 0;
                    const float *xm2 = xc + bx * BSTR;        // box i-2
                    const float *xp1 = xc + (bx + 3) * BSTR;  // box i+1
#pragma unroll 4
                    for (int j = 0; j < b / 4; ++j) {
                        const int t = 4 * j + kr;
                        corr += kr_[t] * xm2[t] - kr_[3 * b + t] * xp1[t];
                    }
                    const float *lm = lsx + bx * LSTR + 2 * gg + c;        // box i-1
                    const float *lb = lsx + (bx + 1) * LSTR + 2 * gg + c;  // box i
                    const float *eh = evs + gg * 2 * TP, *e0 = eh + TP;
#pragma unroll
                    for (int j = 0; j < KT; ++j)
                        corr += eh[4 * j + kr] * lm[(4 * j + kr) * 2 * G] - e0[4 * j + kr] * lb[(4 * j + kr) * 2 * G];
                    corr = lanes_reduce_mod<16>(corr, [](float a, float b2) { return a + b2; });
                    if (kr == 0) acc[0][0] += (2 * sc == P) ? 0.5f * corr : corr;
                }
                // write: pair lanes (c = 0, 1) swap halves (as k_leaf_mma)
#pragma unroll
                for (int rt = 0; rt < NRT; ++rt) {
                    c32 *dst = psec_s + (int64_t)sc * M;
                    const int mu = (i0 + bx) * b + 16 * rt;
                    const float a0 = acc[rt][0], a1 = acc[rt][1], a2 = acc[rt][2], a3 = acc[rt][3];
                    const float keep0 = bsel(c, a2, a0), keep1 = bsel(c, a3, a1);
                    const float recv0 = swap_pair(bsel(c, a0, a2));
                    const float recv1 = swap_pair(bsel(c, a1, a3));
                    *reinterpret_cast<float4 *>(dst + rm(mu + Mfma<float>::row(kr, 2 * c))) =
                        make_float4(bsel(c, recv0, keep0), bsel(c, keep0, recv0), bsel(c, recv1, keep1),
                                    bsel(c, keep1, recv1));
                }
            }
        }
        tile = ntile;
    }
}

// ---------------------------------------------------------------------------
// Tree passes on the matrix cores, one launch per level:
//   mode 0, upward, parent level l (src/fmmfft.c:476-515, mpp :174-184):
//     phi_l[q] = shftfn . phi_{l+1}[2q] + shftfp . phi_{l+1}[2q+1]
//   mode 1, downward, level l >= 3 (src/fmmfft.c:644-749):
//     even c: shftlp . psi_{l-1}[c/2] + flip2p phi_l[c-2] + flip2n phi_l[c+2] + flip3n phi_l[c+3]
//     odd  c: shftln . psi_{l-1}[c/2] + flip3p phi_l[c-3] + flip2p phi_l[c-2] + flip2n phi_l[c+2]
//     node indices periodic (the reference's ring of ranks)
//   mode 2, top flip (src/fmmfft.c:620-639): psi_2[q] = topflp . phi_2[q+2 mod 4]
// Every contraction of a level is out[node][k][row] = sum_t sum_kk
// A_t[k][kk] src_t[node][kk][row] with at most four sources t per node and
// one operator per source that depends only on the node's parity, so for a
// node and a block of 16 rows it is D(16 terms x 16 rows) = sum_t A_t B_t
// with A_t the (zero-padded) operator, read in fragment order (one 256-byte
// load per K step), and B_t read straight from the tree (lane l:
// B[4s + l/16][l%16] = term 4s + l/16 of row 16 rb + l%16).  A wave takes RB
// consecutive row blocks of one node and issues all their loads at once, so
// a level costs about one memory round trip plus its HBM time; consecutive
// waves take the rest of the node, then its neighbours, so the M2L sources
// c +- 2, 3 are shared in one XCD's L2.  TP = terms padded to 16 or 32:
// TP/4 K steps, TP/16 output term tiles.
// ---------------------------------------------------------------------------
// row blocks per wave of the packed form (nodes of at most two row blocks):
// 8 for the FP32 16-term upward pass, else 4 (FP32 downward at P = 16: 4
// measured 1.26 ms per 256 x 2^20 batch against 1.32 at 8 and 1.39 at 2)
template <typename T, int TP>
__host__ __device__ constexpr int tree_mma_rb_packed(int mode) {
    return (mode == 0 && sizeof(T) == 4 && TP == 16) ? 8 : 4;
}

template <typename T, int TP>
__host__ __device__ constexpr int tree_mma_rb(int mode) {
    // row blocks per wave: the B fragments of one wave stay within ~64 VGPRs;
    // fewer where more waves in flight measured faster (FP32 downward, and
    // the 16-term upward pass in both precisions)
    if (mode == 1 && sizeof(T) == 4) return 2;
    if (mode == 0 && TP == 16) return sizeof(T) == 4 ? 1 : 2;
    return mode == 2 ? 4 : (64 * 4 / ((mode == 0 ? 2 : 4) * (TP / 4) * (int)sizeof(T))) >= 4 ? 4
                       : (64 * 4 / ((mode == 0 ? 2 : 4) * (TP / 4) * (int)sizeof(T))) >= 2 ? 2 : 1;
}

// Packed form (PACK, for levels whose nodes have at most two row blocks,
// i.e. P <= 16): a wave takes RB / nrb nodes of one parity (c = 2 RB/nrb
// (w/2) + w%2 + 2k) with all their row blocks, so small nodes do not cost a
// wave (and a memory round trip) each.
//
// SIB4 (downward, complex64, rows a multiple of 64, the large levels): a
// wave takes a sibling pair (c, c + 1), c even, for 64 rows, each lane four
// consecutive rows 64 rg + 4 col + e loaded and stored as one 16-byte access
// (element e of the vector is column col of N tile e).  The pair reads five
// distinct sources instead of eight,
//   c:     shftlp S0 + flip2p S1 + flip2n S3 + flip3n S4
//   c + 1: shftln S0 + flip3p S1 + flip2p S2 + flip2n S4
// with S0 = psi_{l-1}[c/2] and S1..S4 = phi_l[c-2], [c-1], [c+2], [c+3],
// and the 16-byte accesses take a quarter of the address work of the
// one-row-per-lane form (level 14 at N = 2^26: 107 -> 68 us).
// SE (sibling form): rows per lane, 4 (rows a multiple of 64, 16-byte
// accesses) or 2 (32-row nodes, P = 16: 8-byte accesses, 32 rows per wave).
template <typename T, int TP, int MODE, bool PACK = false, bool SPLIT_M2L = false, bool SIB4 = false, int SE = 4>
__global__ __launch_bounds__(256) void k_tree_mma(Geom g, int l, const T *__restrict__ phi_in,
                                                  T *__restrict__ phi_out, const T *__restrict__ psi_in,
                                                  T *__restrict__ psi_out, int64_t tree_stride,
                                                  const T *__restrict__ opfrag, const T *__restrict__ opfrag_lo) {
    using V4 = typename Mfma<T>::v4;
    constexpr int KS = TP / 4, KO = TP / 16;
    constexpr int NT = MODE == 0 ? 2 : MODE == 1 ? 4 : 1;  // sources per node
    // SPLIT_M2L (complex64): the M2L flips (every source of the top flip,
    // sources 1..3 of a downward node) as split products: the FP64 operator
    // is carried as FP32 hi + lo (opfrag_lo = A - fp32(A)), both applied
    constexpr int RB = PACK ? tree_mma_rb_packed<T, TP>(MODE) : tree_mma_rb<T, TP>(MODE);
    constexpr int FR = TP * TP;                             // one operator in fragment order
    constexpr bool SPLIT = SPLIT_M2L && sizeof(T) == 4 && MODE >= 1;
    constexpr int T0 = MODE == 1 ? 1 : 0;                   // first split source
    const int terms = g.terms, rows = tree_rows(g);
    const int node_len = rows * terms;
    const int nn = 1 << l, nrb = (rows + 15) >> 4, nw = (nrb + RB - 1) / RB;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int col = lane & 15, kr = lane >> 4;
    const int wid = 4 * (int)xcd_remap(blockIdx.x, gridDim.x) + wave;
    const int npw = PACK ? RB / nrb : 1;  // nodes per wave
    // first node of the wave and its parity (the same for all its nodes)
    const int c = PACK ? 2 * npw * (wid >> 1) + (wid & 1) : wid / nw;
    const int rb0 = PACK ? 0 : (wid - c * nw) * RB;
    if (c >= nn) return;
    const int64_t ts = (int64_t)blockIdx.z * tree_stride;
    const int odd = c & 1, m = nn - 1;
    // node and row block of item j
    auto item = [&](int j, int &cj, int &rbj) __attribute__((always_inline)) {
        if (PACK) {
            cj = c + 2 * (j / nrb);
            rbj = j - (j / nrb) * nrb;
        } else {
            cj = c;
            rbj = rb0 + j;
        }
    };

    const T *mats[NT];
    if constexpr (MODE == 0) {
        const T *o = opfrag + (size_t)(l - 2) * NOPS * FR;
        mats[0] = o + OP_SHFTFN * FR;
        mats[1] = o + OP_SHFTFP * FR;
    } else if constexpr (MODE == 1) {
        const T *o = opfrag + (size_t)(l - 3) * NOPS * FR;
        mats[0] = o + (odd ? OP_SHFTLN : OP_SHFTLP) * FR;
        mats[1] = o + (odd ? OP_FLIP3P : OP_FLIP2P) * FR;
        mats[2] = o + (odd ? OP_FLIP2P : OP_FLIP2N) * FR;
        mats[3] = o + (odd ? OP_FLIP2N : OP_FLIP3N) * FR;
    } else {
        mats[0] = opfrag + (size_t)(g.n - 2) * NOPS * FR;  // topflp
    }
    // source nodes of node cn (clamped to the level: a packed wave's nodes
    // past the end re-read valid ones and store nothing)
    auto sources = [&](int cn, const T *(&src)[NT]) __attribute__((always_inline)) {
        cn = min(cn, m);
        if constexpr (MODE == 0) {
            const T *ch = phi_in + ts + level_off(g, l + 1);
            src[0] = ch + (int64_t)(2 * cn) * node_len;
            src[1] = ch + (int64_t)(2 * cn + 1) * node_len;
        } else if constexpr (MODE == 1) {
            const T *ph = phi_in + ts + level_off(g, l);
            src[0] = psi_in + ts + level_off(g, l - 1) + (int64_t)(cn >> 1) * node_len;
            src[1] = ph + (int64_t)((cn + (odd ? -3 : -2)) & m) * node_len;
            src[2] = ph + (int64_t)((cn + (odd ? -2 : 2)) & m) * node_len;
            src[3] = ph + (int64_t)((cn + (odd ? 2 : 3)) & m) * node_len;
        } else {
            src[0] = phi_in + ts + level_off(g, 2) + (int64_t)((cn + 2) & 3) * node_len;
        }
    };
    T *lvl_out = (MODE == 0 ? phi_out : psi_out) + ts + level_off(g, l);
    if constexpr (SIB4) {
        static_assert(sizeof(T) == 4 && MODE == 1 && !PACK && TP == 16 && (SE == 4 || SE == 2),
                      "sibling form: complex64, downward");
        using VT = typename std::conditional<SE == 4, float4, float2>::type;
        const int ng = rows / (16 * SE);  // row groups of a node
        const int pr = wid / ng, c0 = 2 * pr;
        if (c0 >= nn) return;
        const int rc = 16 * SE * (wid - pr * ng) + SE * col;
        const T *o = opfrag + (size_t)(l - 3) * NOPS * FR;
        // operators of the even node, then of the odd one (source order above)
        constexpr int opi[8] = {OP_SHFTLP, OP_FLIP2P, OP_FLIP2N, OP_FLIP3N, OP_SHFTLN, OP_FLIP3P, OP_FLIP2P, OP_FLIP2N};
        constexpr int se[4] = {0, 1, 3, 4}, so[4] = {0, 1, 2, 4};
        T a2[8][KS], a2lo[SPLIT ? 8 : 1][KS];
#pragma unroll
        for (int t = 0; t < 8; ++t)
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                a2[t][s] = o[opi[t] * FR + s * 64 + lane];
                if constexpr (SPLIT)
                    if (t % 4 != 0) a2lo[t][s] = (o - opfrag + opfrag_lo)[opi[t] * FR + s * 64 + lane];
            }
        const T *ph = phi_in + ts + level_off(g, l);
        const T *srcs[5] = {psi_in + ts + level_off(g, l - 1) + (int64_t)pr * node_len,
                            ph + (int64_t)((c0 - 2) & m) * node_len, ph + (int64_t)((c0 - 1) & m) * node_len,
                            ph + (int64_t)((c0 + 2) & m) * node_len, ph + (int64_t)((c0 + 3) & m) * node_len};
        // the parent's expansion (source 0) is read by this pair only, and the
        // level's output once by the next level or the leaf: nontemporal hints
        // on both (downward 0.176 -> 0.164 ms at N = 2^26; either alone
        // 0.171 / 0.178); the flips' sources are shared with neighbouring pairs
        typedef float vse __attribute__((ext_vector_type(SE)));
        VT bq[5][KS];
#pragma unroll
        for (int t = 0; t < 5; ++t)
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const T *pa = srcs[t] + min(4 * s + kr, terms - 1) * rows + rc;
                if (t == 0) {
                    const vse q = __builtin_nontemporal_load(reinterpret_cast<const vse *>(pa));
                    if constexpr (SE == 4) bq[t][s] = VT{q.x, q.y, q.z, q.w};
                    else bq[t][s] = VT{q.x, q.y};
                } else {
                    bq[t][s] = *reinterpret_cast<const VT *>(pa);
                }
            }
        V4 acc[2][SE];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int e = 0; e < SE; ++e) acc[h][e] = V4{0, 0, 0, 0};
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int s = 0; s < KS; ++s)
#pragma unroll
                for (int e = 0; e < SE; ++e) {
                    const T be = velem(bq[se[t]][s], e), bo = velem(bq[so[t]][s], e);
                    acc[0][e] = Mfma<T>::mma(a2[t][s], be, acc[0][e]);
                    acc[1][e] = Mfma<T>::mma(a2[4 + t][s], bo, acc[1][e]);
                    if constexpr (SPLIT)
                        if (t != 0) {
                            acc[0][e] = Mfma<T>::mma(a2lo[t][s], be, acc[0][e]);
                            acc[1][e] = Mfma<T>::mma(a2lo[4 + t][s], bo, acc[1][e]);
                        }
                }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            T *dst = lvl_out + (int64_t)(c0 + h) * node_len + rc;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = Mfma<T>::row(kr, r);
                if (k < terms) {
                    vse o;
                    if constexpr (SE == 4) o = vse{acc[h][0][r], acc[h][1][r], acc[h][2][r], acc[h][3][r]};
                    else o = vse{acc[h][0][r], acc[h][1][r]};
                    __builtin_nontemporal_store(o, reinterpret_cast<vse *>(dst + (int64_t)k * rows));
                }
            }
        }
        return;
    }
    // all loads of the wave at once, unconditional (a load under a lane
    // condition becomes a branch that waits for it): operator fragments,
    // then the RB row blocks' source fragments at clamped indices.  No
    // zeroing is needed: rows past the end only feed output columns that are
    // not stored, and terms past the end meet zero operator columns.
    T a[NT][KO][KS], bv[RB][NT][KS], alo[SPLIT ? NT : 1][KO][KS];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int s = 0; s < KS; ++s)
#pragma unroll
            for (int ko = 0; ko < KO; ++ko) {
                a[t][ko][s] = mats[t][(s * KO + ko) * 64 + lane];
                if constexpr (SPLIT)
                    if (t >= T0) alo[t][ko][s] = (mats[t] - opfrag + opfrag_lo)[(s * KO + ko) * 64 + lane];
            }
#pragma unroll
    for (int j = 0; j < RB; ++j) {
        int cj, rbj;
        item(j, cj, rbj);
        const T *src[NT];
        sources(cj, src);
        const int rc = min(16 * rbj + col, rows - 1);
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int s = 0; s < KS; ++s) bv[j][t][s] = src[t][min(4 * s + kr, terms - 1) * rows + rc];
    }
    if constexpr (!PACK) {
        // the RB items' accumulator chains interleaved (independent MFMAs
        // back to back; a single chain waits on each result); items past
        // the level compute on clamped data and store nothing
        V4 acc[RB][KO];
#pragma unroll
        for (int j = 0; j < RB; ++j)
#pragma unroll
            for (int ko = 0; ko < KO; ++ko) acc[j][ko] = V4{0, 0, 0, 0};
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int s = 0; s < KS; ++s)
#pragma unroll
                for (int j = 0; j < RB; ++j)
#pragma unroll
                    for (int ko = 0; ko < KO; ++ko) {
                        acc[j][ko] = Mfma<T>::mma(a[t][ko][s], bv[j][t][s], acc[j][ko]);
                        if constexpr (SPLIT)
                            if (t >= T0) acc[j][ko] = Mfma<T>::mma(alo[t][ko][s], bv[j][t][s], acc[j][ko]);
                    }
#pragma unroll
        for (int j = 0; j < RB; ++j) {
            int cj, rbj;
            item(j, cj, rbj);
            const int row = 16 * rbj + col;
            if (16 * rbj >= rows || cj >= nn) break;  // wave-uniform; items run in order
            T *dst = lvl_out + (int64_t)cj * node_len;
            if (row < rows) {
#pragma unroll
                for (int ko = 0; ko < KO; ++ko)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int k = 16 * ko + Mfma<T>::row(kr, r);
                        if (k < terms) dst[k * rows + row] = acc[j][ko][r];
                    }
            }
        }
    } else {
        // item by item (interleaving all of them made the compiler issue
        // the loads one at a time to save registers)
#pragma unroll
        for (int j = 0; j < RB; ++j) {
            int cj, rbj;
            item(j, cj, rbj);
            const int row = 16 * rbj + col;
            if (16 * rbj >= rows || cj >= nn) break;  // wave-uniform; items run in node order
            T *dst = lvl_out + (int64_t)cj * node_len;
            V4 acc[KO];
#pragma unroll
            for (int ko = 0; ko < KO; ++ko) acc[ko] = V4{0, 0, 0, 0};
#pragma unroll
            for (int t = 0; t < NT; ++t)
#pragma unroll
                for (int s = 0; s < KS; ++s)
#pragma unroll
                    for (int ko = 0; ko < KO; ++ko) {
                        acc[ko] = Mfma<T>::mma(a[t][ko][s], bv[j][t][s], acc[ko]);
                        if constexpr (SPLIT)
                            if (t >= T0) acc[ko] = Mfma<T>::mma(alo[t][ko][s], bv[j][t][s], acc[ko]);
                    }
            if (row < rows) {
#pragma unroll
                for (int ko = 0; ko < KO; ++ko)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int k = 16 * ko + Mfma<T>::row(kr, r);
                        if (k < terms) dst[k * rows + row] = acc[ko][r];
                    }
            }
        }
    }
}

// Two upward levels in one launch (complex64, 16-row-padded terms): the
// wave of (node q of level l-1, row block rb) forms its two children
// phi_l[2q + h] from their children phi_{l+1}[4q + 2h + {0, 1}] and then
// phi_{l-1}[q] from the two accumulators (as B operands in the D layout's K
// order), storing all three (src/fmmfft.c:476-515).
// SUMS: the first P workgroups reduce the section sums instead (no tree
// work).  TOPF (l = 3): the top flip as well, psi_2[(q + 2) mod 4] = topflp
// phi_2[q] (src/fmmfft.c:620-639) from the phi_2 accumulators, which saves
// the top flip its launch.
template <typename T, int TP, bool SUMS = false, bool TOPF = false>
__global__ __launch_bounds__(256) void k_tree_up2(Geom g, int l, T *__restrict__ phi, int64_t tree_stride,
                                                  const T *__restrict__ opfrag, T *__restrict__ psi, SumsJob sj) {
    using V4 = typename Mfma<T>::v4;
    static_assert(TP == 16, "one 16-term block");
    constexpr int KS = TP / 4, FR = TP * TP;
    int bid = blockIdx.x, nb = gridDim.x;
    if constexpr (SUMS) {
        if (bid < g.P) {
            sums_reduce_block(g, sj, bid, blockIdx.z);
            return;
        }
        bid -= g.P;
        nb -= g.P;
    }
    const int terms = g.terms, rows = tree_rows(g);
    const int node_len = rows * terms;
    const int nq = 1 << (l - 1), nrb = (rows + 15) >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int col = lane & 15, kr = lane >> 4;
    const int wid = 4 * (int)xcd_remap(bid, nb) + wave;
    const int q = wid / nrb, rb = wid - q * nrb;
    if (q >= nq) return;
    const int64_t ts = (int64_t)blockIdx.z * tree_stride;
    const int rc = min(16 * rb + col, rows - 1);
    const T *ch = phi + ts + level_off(g, l + 1) + (int64_t)(4 * q) * node_len;
    T *pa = phi + ts + level_off(g, l) + (int64_t)(2 * q) * node_len;
    T *gp = phi + ts + level_off(g, l - 1) + (int64_t)q * node_len;
    const T *o1 = opfrag + (size_t)(l - 2) * NOPS * FR, *o2 = opfrag + (size_t)(l - 3) * NOPS * FR;
    const T *m1[2] = {o1 + OP_SHFTFN * FR, o1 + OP_SHFTFP * FR};
    const T *m2[2] = {o2 + OP_SHFTFN * FR, o2 + OP_SHFTFP * FR};
    T a[2][KS], a2[2][KS], a3[TOPF ? KS : 1], bv[4][KS];
    const T *ot = opfrag + (size_t)(g.n - 2) * NOPS * FR;  // topflp
#pragma unroll
    for (int c2 = 0; c2 < 2; ++c2)
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            a[c2][s] = m1[c2][s * 64 + lane];
            const int j = 16 * (s / 4) + Mfma<T>::row(kr, s % 4);  // D-as-B K order
            a2[c2][s] = m2[c2][(j / 4) * 64 + col + 16 * (j % 4)];
            if constexpr (TOPF)
                if (c2 == 0) a3[s] = ot[(j / 4) * 64 + col + 16 * (j % 4)];
        }
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int s = 0; s < KS; ++s) bv[k][s] = ch[(int64_t)k * node_len + min(4 * s + kr, terms - 1) * rows + rc];
    __builtin_amdgcn_sched_barrier(0);
    V4 up[2] = {V4{0, 0, 0, 0}, V4{0, 0, 0, 0}};
#pragma unroll
    for (int c2 = 0; c2 < 2; ++c2)
#pragma unroll
        for (int s = 0; s < KS; ++s)
#pragma unroll
            for (int h = 0; h < 2; ++h) up[h] = Mfma<T>::mma(a[c2][s], bv[2 * h + c2][s], up[h]);
    V4 g0 = V4{0, 0, 0, 0}, g1 = V4{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        g0 = Mfma<T>::mma(a2[0][s], up[0][s % 4], g0);
        g1 = Mfma<T>::mma(a2[1][s], up[1][s % 4], g1);
    }
    const V4 gv = g0 + g1;
    V4 pt = V4{0, 0, 0, 0};
    if constexpr (TOPF)
#pragma unroll
        for (int s = 0; s < KS; ++s) pt = Mfma<T>::mma(a3[s], gv[s % 4], pt);
    const int row = 16 * rb + col;
    if (row < rows) {
        auto store = [&](T *dst, const V4 &v) __attribute__((always_inline)) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = Mfma<T>::row(kr, r);
                if (k < terms) dst[k * rows + row] = v[r];
            }
        };
        store(pa, up[0]);
        store(pa + node_len, up[1]);
        store(gp, gv);
        if constexpr (TOPF) store(psi + ts + level_off(g, 2) + (int64_t)((q + 2) & 3) * node_len, pt);
    }
}

// Two downward levels in one launch (complex64, 16-row-padded terms): the
// wave of (node c of level l, row block rb) forms psi_l[c] as mode 1 does,
// stores it, and then forms both children psi_{l+1}[2c + h] for the same
// rows: their L2L term takes psi_l[c] straight from the accumulators (as B
// operands in the D layout's K order, the operator fragments read in that
// order) and their three flips read phi_{l+1} as usual.  Saves the launch
// and the re-read of psi_l of the per-level form (src/fmmfft.c:644-749).
template <typename T, int TP, bool SPLIT = false>
__global__ __launch_bounds__(256) void k_tree_down2(Geom g, int l, const T *__restrict__ phi, T *__restrict__ psi,
                                                    int64_t tree_stride, const T *__restrict__ opfrag,
                                                    const T *__restrict__ opfrag_lo) {
    using V4 = typename Mfma<T>::v4;
    static_assert(TP == 16, "one 16-term block");
    constexpr int KS = TP / 4, FR = TP * TP;
    const int terms = g.terms, rows = tree_rows(g);
    const int node_len = rows * terms;
    const int nn = 1 << l, nrb = (rows + 15) >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int col = lane & 15, kr = lane >> 4;
    const int wid = 4 * (int)xcd_remap(blockIdx.x, gridDim.x) + wave;
    const int c = wid / nrb, rb = wid - c * nrb;
    if (c >= nn) return;
    const int64_t ts = (int64_t)blockIdx.z * tree_stride;
    const int rc = min(16 * rb + col, rows - 1);
    const int odd = c & 1;
    // operators of a node of level lv and parity par: L2L, then the three
    // flips in the order of flip_node's sources 1..3
    auto op = [&](int lv, int par, int t) -> const T * {
        const T *o = opfrag + (size_t)(lv - 3) * NOPS * FR;
        const int id = t == 0 ? (par ? OP_SHFTLN : OP_SHFTLP)
                     : t == 1 ? (par ? OP_FLIP3P : OP_FLIP2P)
                     : t == 2 ? (par ? OP_FLIP2P : OP_FLIP2N)
                              : (par ? OP_FLIP2N : OP_FLIP3N);
        return o + id * FR;
    };
    auto flip_node = [&](int lv, int cn, int par, int t) -> int64_t {
        const int off = t == 1 ? (par ? -3 : -2) : t == 2 ? (par ? -2 : 2) : (par ? 2 : 3);
        return (int64_t)((cn + off) & ((1 << lv) - 1)) * node_len;
    };
    const T *phl = phi + ts + level_off(g, l), *phc = phi + ts + level_off(g, l + 1);
    T *psl = psi + ts + level_off(g, l), *psc = psi + ts + level_off(g, l + 1);
    const T *psp = psi + ts + level_off(g, l - 1) + (int64_t)(c >> 1) * node_len;

    // every load of the wave at once, unconditional at clamped indices; with
    // SPLIT the flips (M2L) carry their FP32 residual operators too
    T a[4][KS], bv[4][KS], al[2][KS], af[2][3][KS], bc[2][3][KS], alo[4][KS], aflo[2][3][KS];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const T *m = op(l, odd, t);
        const T *src = t == 0 ? psp : phl + flip_node(l, c, odd, t);
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            a[t][s] = m[s * 64 + lane];
            if constexpr (SPLIT) alo[t][s] = t == 0 ? T(0) : (m - opfrag + opfrag_lo)[s * 64 + lane];
            bv[t][s] = src[min(4 * s + kr, terms - 1) * rows + rc];
        }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const T *m0 = op(l + 1, h, 0);
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            // K step s of the D-as-B product: lane group kr holds term j
            const int j = 16 * (s / 4) + Mfma<T>::row(kr, s % 4);
            al[h][s] = m0[(j / 4) * 64 + col + 16 * (j % 4)];
        }
#pragma unroll
        for (int t = 1; t < 4; ++t) {
            const T *m = op(l + 1, h, t);
            const T *src = phc + flip_node(l + 1, 2 * c + h, h, t);
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                af[h][t - 1][s] = m[s * 64 + lane];
                if constexpr (SPLIT) aflo[h][t - 1][s] = (m - opfrag + opfrag_lo)[s * 64 + lane];
                bc[h][t - 1][s] = src[min(4 * s + kr, terms - 1) * rows + rc];
            }
        }
    }
    // (the loads stay ahead of every MFMA: 0.2336 -> 0.2316 ms downward)
    __builtin_amdgcn_sched_barrier(0);
    // four independent chains: psi_l in two halves, the children's flips
    V4 p0 = V4{0, 0, 0, 0}, p1 = V4{0, 0, 0, 0}, ch[2] = {V4{0, 0, 0, 0}, V4{0, 0, 0, 0}};
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            if (t < 2) {
                p0 = Mfma<T>::mma(a[2 * t][s], bv[2 * t][s], p0);
                p1 = Mfma<T>::mma(a[2 * t + 1][s], bv[2 * t + 1][s], p1);
                if constexpr (SPLIT) {
                    if (t == 1) p0 = Mfma<T>::mma(alo[2][s], bv[2][s], p0);
                    p1 = Mfma<T>::mma(alo[2 * t + 1][s], bv[2 * t + 1][s], p1);
                }
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                ch[h] = Mfma<T>::mma(af[h][t][s], bc[h][t][s], ch[h]);
                if constexpr (SPLIT) ch[h] = Mfma<T>::mma(aflo[h][t][s], bc[h][t][s], ch[h]);
            }
        }
    const V4 pl = p0 + p1;
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
        for (int h = 0; h < 2; ++h) ch[h] = Mfma<T>::mma(al[h][s], pl[s % 4], ch[h]);
    const int row = 16 * rb + col;
    if (row < rows) {
        auto store = [&](T *dst, const V4 &v) __attribute__((always_inline)) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = Mfma<T>::row(kr, r);
                if (k < terms) dst[k * rows + row] = v[r];
            }
        };
        store(psl + (int64_t)c * node_len, pl);
        store(psc + (int64_t)(2 * c) * node_len, ch[0]);
        store(psc + (int64_t)(2 * c + 1) * node_len, ch[1]);
    }
}

// ---------------------------------------------------------------------------
// Combine (src/fmmfft.c:956-970): for every row mu,
//   y_k2[mu] = x_0[mu] + sum_sc w_P^{sc k2} gfac_sc (psec[sc][mu] - sumsec_sc),
// i.e. the reference's dense fo product done as a P-point FFT (Stockham in
// LDS, FP64 arithmetic for both precisions) and written block-major (the
// reference's MPI_Alltoall).  Tiles of 32 rows; reads and writes are 32
// contiguous elements per section / block.  Persistent workgroups prefetch
// the next tile into registers while transforming the current one.
// Rows are independent, so the kernel is blind to the row order: psec and y
// share whatever layout the leaf kernel wrote (natural or column-tiled).
// ---------------------------------------------------------------------------
constexpr int kCombineThreads = 256;

// Rows mu per combine tile: 2048 elements per tile (16 KB of complex64
// read) for P >= 32, 1024 for smaller P (more workgroups in flight measured
// faster at P = 16: 1.04 -> 0.98 ms for 256 signals of 2^20).
template <int P>
constexpr int combine_rows() { return P >= 32 ? 2048 / P : 1024 / P; }

template <typename T, int P, bool TAB = false>
__global__ __launch_bounds__(kCombineThreads) void k_combine(Geom g, const cx<T> *__restrict__ psec, int64_t psec_stride,
                                                 const c64 *__restrict__ sumsec, int64_t sumsec_stride,
                                                 const c64 *__restrict__ gfac, const c64 *__restrict__ wP,
                                                 SigAddr<cx<T>> ya) {
    constexpr int RB = combine_rows<P>();
    constexpr int NT = kCombineThreads;
    // psec is dead after this read and y is read once by the next pass:
    // nontemporal loads and stores (N = 2^26: 0.203 -> 0.198 ms complex64,
    // 0.446 -> 0.432 ms complex128)
    constexpr bool CMB_NT_LD = true, CMB_NT_ST = true;
    // one pad slot per 8 elements (an 8-way bank conflict in the first
    // radix-8 stage's writes with one per 16: combine 0.224 -> 0.210 ms)
    constexpr int ls = P + P / 8 + 1;
    constexpr int PER = (RB * P + NT - 1) / NT;
    __shared__ c64 a[RB * ls];
    const int sig = blockIdx.z;
    psec += (int64_t)sig * psec_stride;
    sumsec += (int64_t)sig * sumsec_stride;
    cx<T> *__restrict__ y = ya.template at<TAB>(sig);
    const int64_t ntiles = (g.M + RB - 1) / RB;
    const int64_t nwg = gridDim.x;
    const int64_t xo = xcd_remap(blockIdx.x, nwg);

    // per-lane section constants (a lane's sections are fixed across tiles)
    c64 gf[PER], ss[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int e = threadIdx.x + u * NT;
        const int sc = e / RB;
        gf[u] = (e < RB * P && sc != 0) ? gfac[sc] : c64{1.0, 0.0};
        ss[u] = (e < RB * P && sc != 0) ? sumsec[sc] : c64{0.0, 0.0};
    }
    // prefetch in scalar registers; a ragged last tile re-reads its last
    // valid row (unconditional loads keep the arrays out of scratch)
    T vr[PER], vi[PER];
    auto prefetch = [&](int64_t t) __attribute__((always_inline)) {
        const int64_t mu0 = t * RB;
        const int rl = (int)min((int64_t)RB, g.M - mu0) - 1;
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int e = threadIdx.x + u * NT;
            const int sc = min(e / RB, P - 1), r = min(e % RB, rl);
            const cx<T> x = CMB_NT_LD ? ld_nt(psec + (int64_t)sc * g.M + mu0 + r) : psec[(int64_t)sc * g.M + mu0 + r];
            vr[u] = x.re;
            vi[u] = x.im;
        }
    };
    int64_t it = 0, t = xo;
    if (t < ntiles) prefetch(t);
    while (t < ntiles) {
        __syncthreads();  // the previous tile's outputs have left LDS
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int e = threadIdx.x + u * NT;
            const int sc = e / RB, r = e - sc * RB;
            if (e >= RB * P) continue;
            const c64 w{(double)vr[u] - ss[u].re, (double)vi[u] - ss[u].im};
            a[r * ls + lpadp<3>(sc)] = cmul(gf[u], w);
        }
        __syncthreads();
        const int64_t tn = (++it) * nwg + xo;
        if (tn < ntiles) prefetch(tn);
        stockham<double, P, RB, 1, NT, 3>(a, ls, wP);
        const int64_t mu0 = t * RB;
        const int rv = (int)min((int64_t)RB, g.M - mu0);
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int e = threadIdx.x + u * NT;
            const int k2 = e / RB, r = e - k2 * RB;
            if (e >= RB * P || r >= rv) continue;
            const c64 o = a[r * ls + lpadp<3>(k2)];
            if (CMB_NT_ST) st_nt(y + (int64_t)k2 * g.M + mu0 + r, cx<T>{(T)o.re, (T)o.im});
            else y[(int64_t)k2 * g.M + mu0 + r] = cx<T>{(T)o.re, (T)o.im};
        }
        t = tn;
    }
}

}  // namespace fmmfft
