// Shared device types and helpers of the MI355X FMM-FFT kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace fmmfft {

// Complex value in the plan's storage precision; 8- or 16-byte aligned so
// that a load is one dwordx2 / dwordx4.
template <typename T>
struct alignas(2 * sizeof(T)) cx {
    T re, im;
};

using c32 = cx<float>;
using c64 = cx<double>;

// Slots of the per-level translation operators (the reference's shftfn,
// shftfp, flip2n, flip2p, flip3n, flip3p, shftln, shftlp,
// src/fmmfft.c:281-290).  Level slot L holds the M2M pair with a child at
// level L+3, the flips of level L+3 and the L2L pair into a child at level
// L+3.
enum : int {
    OP_SHFTFN = 0,
    OP_SHFTFP = 1,
    OP_FLIP2N = 2,
    OP_FLIP2P = 3,
    OP_FLIP3N = 4,
    OP_FLIP3P = 5,
    OP_SHFTLN = 6,
    OP_SHFTLP = 7,
    NOPS = 8
};

// Operator tables in one precision, all device pointers.
template <typename T>
struct Tables {
    const T *initch;  // [b][terms]
    const T *ops;     // [n-2][8][terms][terms]
    const T *topflp;  // [terms][terms]
    const T *evalm;   // [P-1][b][terms]
    const T *evalmh;  // [P/2][terms]
    const T *kv;      // [P-1][4b]   K_sc(d), d = -2b .. 2b-1
};

// Geometry of one transform; passed by value to every kernel.
struct Geom {
    int64_t N, M;   // N = P * M
    int P, S;       // S = P - 1 sections carry a cot matvec
    int terms, b;   // Chebyshev order, points per leaf box
    int s, n;       // leaf boxes per section (2^n), tree depth
};

// Signal sig of a batch in a caller's buffer: base + sig * stride
// elements, or, for a batch of separate buffers (tab set), tab[sig], a
// device table of their addresses.  The kernels that read or write callers'
// buffers take a template flag TAB: without it they use the contiguous form
// only and take no table load (a per-tile table load, behind a branch or
// not, made their counter waits cover other loads in flight: +4..16 us per
// kernel at N = 2^26).
template <typename E>
struct SigAddr {
    E *base;
    int64_t stride;
    E *const *tab;
    template <bool TAB>
    __device__ inline E *at(int64_t sig) const {
        if constexpr (TAB)
            if (tab) return tab[sig];
        return base + sig * stride;
    }
};

// Loads and stores of one complex point with the nontemporal hint (data
// streamed once: not kept in the caches for reuse).
template <typename T>
__device__ inline cx<T> ld_nt(const cx<T> *p) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 t = __builtin_nontemporal_load(reinterpret_cast<const v2 *>(p));
    return cx<T>{t.x, t.y};
}
template <typename T>
__device__ inline void st_nt(cx<T> *p, cx<T> v) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    __builtin_nontemporal_store(v2{v.re, v.im}, reinterpret_cast<v2 *>(p));
}

// Row order of the leaf output (psec) and of the combine output.  Natural
// (ltw < 0), or column-tiled for a two-pass local FFT of M = M1 x M2
// (M1 = 2^lm1, M2 = 2^lm2, row mu = n1 M2 + n2): row mu at
// (n2 / TW) TW M1 + n1 TW + n2 % TW, TW = 2^ltw the column pass's tile
// width, so that every column-pass tile (TW columns of M1 points) is one
// contiguous range.  Offsets inside a block fit 32 bits.
struct RowMap {
    int ltw, lm1, lm2;
    __host__ __device__ inline int operator()(int mu) const {
        if (ltw < 0) return mu;
        const int n1 = mu >> lm2, n2 = mu & ((1 << lm2) - 1);
        return ((n2 >> ltw) << (ltw + lm1)) + (n1 << ltw) + (n2 & ((1 << ltw) - 1));
    }
};

// L2T K steps (of 4 terms) per row tile of the matrix-core leaf kernel; the
// evaluation-row table and the kernel's LDS copy of the local expansions are
// padded to 4 * leaf_kt(terms) terms.
__host__ __device__ inline int leaf_kt(int terms) { return terms <= 16 ? 4 : 8; }

// Tree nodes are [terms][2P] reals: row 2 sc + re/im of section sc.
// Section 0 has no cot matvec, but its rows are kept (they carry its
// expansion, which nothing reads): a term row is then 2P reals, a multiple
// of 64 bytes for P >= 8, and a block of 16 rows is 8 whole sections whose
// reals sit at the same offsets as in a row of x.
__host__ __device__ inline int tree_rows(const Geom &g) { return 2 * g.P; }
// Real offset of level l (levels 2..n, 2^l nodes).
__host__ __device__ inline int64_t level_off(const Geom &g, int l) {
    return ((int64_t(1) << l) - 4) * int64_t(tree_rows(g)) * g.terms;
}
__host__ __device__ inline int64_t tree_len(const Geom &g) {
    return ((int64_t(1) << (g.n + 1)) - 4) * int64_t(tree_rows(g)) * g.terms;
}

__device__ inline int pmod(int a, int m) {
    int r = a % m;
    return r < 0 ? r + m : r;
}

__device__ inline c64 cmul(c64 a, c64 b) {
    return c64{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ inline cx<T> cmulT(cx<T> a, cx<T> b) {
    return cx<T>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

}  // namespace fmmfft
