// C ABI of the MI355X FMM-FFT: plans, device tables, workspace and the
// execute launch sequence (include/fmmfft.h).
#include "../../include/fmmfft.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <new>

#include "common.hpp"
#include "internal.hpp"
#include "kernels_fft.hpp"
#include "kernels_fmm.hpp"
#ifndef FMMFFT_TAB_UNIT
#include "kernels_ops.hpp"  // table construction: plan creation only
#endif

// This file is compiled twice: as itself (plans, the C ABI, executes on
// contiguous batches) and, with FMMFFT_TAB_UNIT defined, by fmmfft_tab.hip
// for the table-addressed executes of the server (fmmfft::execute_table),
// so that the two sets of kernel instances build in parallel.
using namespace fmmfft;

#ifndef FMMFFT_TAB_UNIT
std::atomic<int> fmmfft::g_last_hip_error{0};
#endif

namespace {

#ifndef FMMFFT_TAB_UNIT
// Workspace of fmmfft_execute_static: one array per device, held for the
// life of the process.  Host code never uses this name as an address (on
// the host it is only a stand-in): the device address comes from
// hipGetSymbolAddress, once per device, in static_scratch().
static __device__ __align__(256) unsigned char scratch[FMMFFT_STATIC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch (nullptr on failure; the
// HIP error is recorded).
void *static_scratch(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            g_last_hip_error.store((int)e);
            return nullptr;
        }
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}
#endif

constexpr size_t kAlign = 256;
inline size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }
inline int ilog2(int64_t v) {
    int r = 0;
    while (v > 1) { v >>= 1; ++r; }
    return r;
}
inline bool is_pow2(int64_t v) { return v > 0 && (v & (v - 1)) == 0; }

#define HIP_TRY(expr)                                          \
    do {                                                       \
        hipError_t e_ = (expr);                                \
        if (e_ != hipSuccess) {                                \
            g_last_hip_error.store((int)e_);                   \
            return e_ == hipErrorOutOfMemory ? FMMFFT_ENOMEM   \
                                             : FMMFFT_EHIP;    \
        }                                                      \
    } while (0)

struct Params {
    Geom g;
    int dtype;
    size_t esize;  // bytes of one real in the plan's precision
};

// complex64 storage (plain, or with split-product M2L flips)
inline bool is_c64(int dtype) { return dtype == FMMFFT_C64 || dtype == FMMFFT_C64_SPLIT; }

int validate(int64_t n, int32_t P, int32_t terms, int32_t leaf, int32_t dtype, Params *out) {
    if (!is_c64(dtype) && dtype != FMMFFT_C128) return FMMFFT_EINVAL;
    if (P < 2 || P > 64 || !is_pow2(P) || n <= 0 || n % P) return FMMFFT_EINVAL;
    if (terms < 2 || terms > 32 || leaf < 2 || leaf > 128) return FMMFFT_EINVAL;
    const int64_t M = n / P;
    if (M % leaf) return FMMFFT_EINVAL;
    const int64_t s = M / leaf;
    if (!is_pow2(s) || s < 8) return FMMFFT_EINVAL;
    if (leaf % 4) return FMMFFT_ENOTSUP;                   // leaf lanes own 4 targets
    if ((int64_t)leaf * P > (is_c64(dtype) ? 8192 : 4096)) return FMMFFT_ENOTSUP;
    // local FFTs in at most three passes of <= 1024 points; offsets inside
    // one signal fit 32 bits
    if (!is_pow2(M) || M > (int64_t(1) << 30) || n > (int64_t(1) << 30)) return FMMFFT_ENOTSUP;
    Geom &g = out->g;
    g.N = n;
    g.M = M;
    g.P = P;
    g.S = P - 1;
    g.terms = terms;
    g.b = leaf;
    g.s = (int)s;
    g.n = ilog2(s);
    out->dtype = dtype;
    out->esize = is_c64(dtype) ? 4 : 8;
    return FMMFFT_OK;
}

// Offsets (in doubles) of the FP64 operator tables; the FP32 copies use the
// same offsets in floats for the first `n_ops` entries.
struct TableLayout {
    size_t cheb, initch, ops, topflp, evalm, evalmh, kv, evfrag, opfrag, n_ops;  // contiguous block
    size_t gfac, wP;                                             // complex128 [P]
    size_t twA, twB, twC, tlo, thi;                              // complex128
    size_t total;                                                // doubles
};

// Local FFT of length M in 1, 2 or 3 passes of sub-FFTs of length <= 1024:
// M = M1 (one pass), M1 M2 (M <= 2^20) or M1 M2 M3 (M <= 2^30; the first
// pass transforms along the M1 axis with the rest, R = M2 M3, as columns).
struct FftLayout {
    bool two_pass;    // more than one pass
    int passes;
    int64_t M1, M2, M3;
    int logR;
};

FftLayout fft_layout(const Geom &g) {
    FftLayout f{};
    const int m = ilog2(g.M);
    f.M2 = f.M3 = 1;
    if (g.M <= 1024) {
        f.passes = 1;
        f.M1 = g.M;
    } else if (m <= 20) {
        f.passes = 2;
        f.M1 = int64_t(1) << ((m + 1) / 2);
        f.M2 = int64_t(1) << (m / 2);
    } else {
        // m = 21..30: three factors of 2^7..2^10 (the last >= the column
        // tile width of the middle pass)
        f.passes = 3;
        const int m1 = (m + 2) / 3, m2 = (m - m1 + 1) / 2, m3 = m - m1 - m2;
        f.M1 = int64_t(1) << m1;
        f.M2 = int64_t(1) << m2;
        f.M3 = int64_t(1) << m3;
    }
    f.two_pass = f.passes > 1;
    f.logR = (m + 1) / 2;
    return f;
}

TableLayout table_layout(const Geom &g, const FftLayout &f) {
    TableLayout t{};
    size_t o = 0;
    const size_t T = g.terms;
    t.cheb = o; o += T;
    t.initch = o; o += (size_t)g.b * T;
    t.ops = o; o += (size_t)(g.n - 2) * NOPS * T * T;
    t.topflp = o; o += T * T;
    t.evalm = o; o += (size_t)g.S * g.b * T;
    t.evalmh = o; o += (size_t)(g.P / 2) * T;
    t.kv = o; o += (size_t)g.S * 4 * g.b;
    t.evfrag = o; o += (g.b % 16 == 0) ? (size_t)g.S * (g.b / 16) * leaf_kt(g.terms) * 64 : 0;
    const size_t tp = 4 * (size_t)leaf_kt(g.terms);
    t.opfrag = o; o += ((size_t)(g.n - 2) * NOPS + 1) * tp * tp;
    t.n_ops = o;
    o = (o + 1) & ~size_t(1);
    t.gfac = o; o += 2 * (size_t)g.P;
    t.wP = o; o += 2 * (size_t)g.P;
    t.twA = o; o += 2 * (size_t)f.M1;  // full circle: e^{-2 pi i k / M1}, k < M1
    t.twB = o; o += 2 * (size_t)f.M2;
    t.twC = o; o += 2 * (size_t)f.M3;
    const int64_t R = int64_t(1) << f.logR;
    t.tlo = o; o += 2 * (size_t)R;
    t.thi = o; o += 2 * (size_t)std::max<int64_t>(1, g.M / R);
    t.total = o;
    return t;
}

// Launch shapes of the leaf and S2M kernels.
struct Shapes {
    int s2m_bpb, s2m_nblk;
    int s2m_levels = 0;  // tree levels above the leaves the S2M kernel forms
    size_t s2m_lds;
    int s2m_vec;  // matrix-core S2M: reals per lane load (0: vector-FMA kernel)
    RowMap rm;    // row order of the leaf output (column-tiled for two FFT passes)
    int leaf_G, leaf_bb;
    size_t leaf_lds;
    bool leaf_mma;  // matrix-core leaf kernel (leaf size a multiple of 16)
    bool leaf_h16;  // complex64 leaf 32 / 64: near field as FP16 split products
};

// Register/LDS padding of the Chebyshev order for the row kernels.
inline int padded_terms(int terms) { return terms <= 8 ? 8 : terms <= 16 ? 16 : terms <= 24 ? 24 : 32; }

size_t leaf_lds(const Geom &g, size_t esize, int G, int bb) {
    return (size_t)(bb + 3) * g.b * G * 2 * esize + (size_t)G * (4 * g.b + 1) * esize +
           (size_t)(bb + 1) * G * 2 * g.terms * esize;
}

// Column tiles (of 8 boxes) per leaf tile: leaf 16 takes 4 so that a tile
// holds as many MFMAs as one of leaf 64 (at most s / 8); leaf 32 measured
// faster with one column tile at four waves per SIMD than with two at two
// (batched 256 x 2^20 leaf stage 2.32 -> 2.17 ms).
int leaf_ct(const Geom &g) {
    const int ct = g.b >= 32 ? 1 : 4;
    return std::min(ct, g.s / 8);
}

size_t leaf_h16_lds(const Geom &g, int G) {
    const int nrt = g.b / 16;
    const size_t wstr = nrt == 4 ? (size_t)LeafH16<4>::WSTR : (size_t)LeafH16<2>::WSTR;
    const size_t tp = 4 * (size_t)leaf_kt(g.terms), lstr = tp * 2 * G + 2;
    const size_t nwv = G > 4 ? G : 4;
    return (2 * G * wstr + (size_t)G * 4 * g.b + 9 * lstr + (size_t)G * 2 * tp + nwv * (size_t)G) * sizeof(float);
}

// sections per tile of the FP16 leaf kernel (one per wave).  Leaf 32 takes
// 8 (64-byte pieces of each x row; one 8-wave workgroup per CU at 158
// VGPRs): 256 x 2^20 batch at P = 16, leaf 1.94 -> 1.69 ms against 4
// sections (32-byte pieces, three 4-wave workgroups per CU); 16 sections
// (whole 128-byte rows) need 1024 threads at 128 VGPRs, spill and took
// 2.13 ms.  Leaf 64 (236 VGPRs) stays at 4 (8: 0.438 against 0.432 ms at
// N = 2^26).
inline int leaf_h16_G(const Geom &g) { return std::min(g.P, g.b == 32 ? 8 : 4); }

size_t leaf_mma_lds(const Geom &g, size_t esize, int G) {
    const int bb = 8 * leaf_ct(g);
    const int q = bb + 2, xpad = (q & 3) + 8 * (q >> 2) + 1;
    const size_t tp = 4 * (size_t)leaf_kt(g.terms);
    return ((size_t)(bb + 3) * g.b * G + xpad) * 2 * esize + (size_t)G * 4 * g.b * esize +
           (size_t)(bb + 1) * (tp * 2 * G + 2) * esize + (size_t)G * 2 * tp * esize;
}
// (the persistent leaf kernel's LDS is at most 11*128*4*16 + ... < 100 KB)

Shapes shapes(const Geom &g, size_t esize) {
    Shapes sh{};
    const size_t lds_cap = 64 * 1024;
    // S2M: tiles of 64 KB of x (bx boxes of all sections), at most 512
    // persistent workgroups (one partial section sum each)
    const int s2m_bx = std::min<int>((int)(65536 / (2 * esize) / ((size_t)g.b * g.P)), g.s);
    sh.s2m_bpb = s2m_bx;
    sh.s2m_nblk = std::min(g.s / s2m_bx, 512);
    sh.s2m_lds = (size_t)g.b * padded_terms(g.terms) * esize;
    // matrix-core S2M for complex64, leaf 16..64: a row of x (2P reals)
    // splits into chunks of 16 VEC reals, VEC = 4 (P >= 32) or 2 (P = 16);
    // one chunk per wave, 4 / chunks sibling box pairs per workgroup step
    // (the kernel also computes their parents, level n-1), at most 768
    // workgroups (three per CU: the leaf-64 instances hold 3 to 5 per CU).
    // complex128 keeps the vector-FMA kernel, which already streams x at
    // copy speed (the FP64 VALU rate equals the FP32 one here).
    {
        int vec = 0;
        if (esize == 4 && g.b % 16 == 0 && g.b <= 64) {
            if (2 * g.P >= 64) vec = 4;
            else if (2 * g.P >= 32) vec = 2;
        }
        sh.s2m_vec = vec;
        if (vec) {
            // sibling quads (levels n-1 and n-2 formed too) when the tree
            // has them and the terms fit one 16-row block (with 32 the quad
            // form holds one wave per SIMD), else sibling pairs
            const int nchunk = 2 * g.P / (16 * vec), pper = 4 / nchunk;
            sh.s2m_levels = (g.n >= 4 && g.terms <= 16) ? 2 : 1;
            const int units = g.s >> sh.s2m_levels;
            sh.s2m_nblk = std::min((units + pper - 1) / pper, 768);
        }
    }

    int G = std::min(g.P, esize == 4 ? 16 : 8);
    int bb = std::min(g.s, 4);
    while (leaf_lds(g, esize, G, bb) > lds_cap && bb > 1) bb >>= 1;
    while (leaf_lds(g, esize, G, bb) > lds_cap && G > 1) G >>= 1;
    sh.leaf_G = G;
    sh.leaf_bb = bb;
    sh.leaf_lds = leaf_lds(g, esize, G, bb);
    sh.leaf_mma = g.b % 16 == 0 && g.b <= 128;  // b is a power of two: b/16 in {1,2,4,8}
    // two-pass local FFT: the leaf writes psec column-tiled (the combine
    // keeps the order) in tiles of TW = max(C, 128 bytes) columns, C the
    // column pass's tile width: the leaf keeps 128-byte output pieces and a
    // column-pass tile reads C-column pieces at a stride of TW (one 128-byte
    // line per row for a pair of neighbouring tiles) instead of M2
    {
        const FftLayout fl = fft_layout(g);
        if (fl.two_pass) {
            const int cols = esize == 4 ? fft_cols<float>((int)fl.M1) : fft_cols<double>((int)fl.M1);
            const int tw = std::max(cols, (int)(128 / (2 * esize)));
            // (tw <= 32 <= M / M1 whenever several passes run; kept as a guard)
            const int64_t R = g.M / fl.M1;
            sh.rm = tw <= R ? RowMap{ilog2(tw), ilog2(fl.M1), ilog2(R)} : RowMap{-1, 0, 0};
        } else {
            sh.rm = RowMap{-1, 0, 0};
        }
    }
    if (sh.leaf_mma) {
        // one 64-byte x segment per window row, 8 boxes per workgroup
        // 4 sections per tile (2 when P = 2): one section per wave
        int Gm = std::min(g.P, 4);
        sh.leaf_G = Gm;
        sh.leaf_bb = 8;
        sh.leaf_lds = leaf_mma_lds(g, esize, Gm);
        sh.leaf_h16 = esize == 4 && (g.b == 32 || g.b == 64);
        if (sh.leaf_h16) {
            sh.leaf_G = leaf_h16_G(g);
            sh.leaf_lds = leaf_h16_lds(g, sh.leaf_G);
        }
    }


    return sh;
}

struct WsLayout {
    size_t sums_part, sumsec, phi, psi, psec, stride;
};

WsLayout ws_layout(const Geom &g, size_t esize, const Shapes &sh) {
    WsLayout w{};
    size_t o = 0;
    w.sums_part = o; o = align_up(o + (size_t)sh.s2m_nblk * g.P * 16);
    w.sumsec = o; o = align_up(o + (size_t)(g.P + 1) * 16);
    const size_t tl = (size_t)tree_len(g) * esize;
    w.phi = o; o = align_up(o + tl);
    // the leaf reads the finest local expansions beside the rows of x (both
    // at the same offsets inside 2P-real rows): starting them at an odd
    // multiple of 256 bytes puts their requests on the other half of the
    // 256-byte-interleaved memory channels (N = 2^26, P = 64: leaf 0.437 ->
    // 0.427 ms against a multiple of 512)
    if ((o + (size_t)level_off(g, g.n) * esize) % 512 < 256) o += 256;
    w.psi = o; o = align_up(o + tl);
    w.psec = o; o = align_up(o + (size_t)g.N * 2 * esize);
    w.stride = o;
    return w;
}

}  // namespace

struct fmmfft_plan_s {
    Params prm;
    FftLayout fft;
    TableLayout tl;
    Shapes sh;
    WsLayout ws;
    int device;
    double *tab64 = nullptr;  // FP64 tables (TableLayout offsets)
    float *tab32 = nullptr;   // FP32 copies of the operator block + twiddles
    _Float16 *kfrag = nullptr;  // FP16 split near-field fragments (leaf_h16)
    size_t tw32_A = 0, tw32_B = 0, tw32_C = 0;  // float offsets of FP32 stage twiddles
    size_t opf_lo32 = 0;            // float offset of the FP32 residual operator fragments
    void *work = nullptr;     // default workspace
    int64_t max_batch = 1;
    // profiling: FMMFFT_NSTAGES + 1 events per recorded execute
    hipEvent_t *prof_ev = nullptr;
    int32_t prof_cap = 0;
    int64_t prof_calls = 0;
};

namespace {

template <typename T>
struct TabPtrs {
    const T *initch, *ops, *topflp, *evalm, *evalmh, *kv, *evfrag, *opfrag, *twA, *twB, *twC;
    const T *opfrag_lo;  // FP32 residual of the fragment-ordered operators (complex64), else null
};

template <typename T>
TabPtrs<T> tab_ptrs(const fmmfft_plan_s *p);

template <>
TabPtrs<double> tab_ptrs<double>(const fmmfft_plan_s *p) {
    const double *t = p->tab64;
    return {t + p->tl.initch, t + p->tl.ops, t + p->tl.topflp, t + p->tl.evalm, t + p->tl.evalmh,
            t + p->tl.kv,     t + p->tl.evfrag, t + p->tl.opfrag, t + p->tl.twA, t + p->tl.twB, t + p->tl.twC,
            nullptr};
}
template <>
TabPtrs<float> tab_ptrs<float>(const fmmfft_plan_s *p) {
    const float *t = p->tab32;
    return {t + p->tl.initch, t + p->tl.ops, t + p->tl.topflp, t + p->tl.evalm, t + p->tl.evalmh,
            t + p->tl.kv,     t + p->tl.evfrag, t + p->tl.opfrag, t + p->tw32_A, t + p->tw32_B, t + p->tw32_C,
            p->prm.dtype == FMMFFT_C64_SPLIT ? t + p->opf_lo32 : nullptr};
}

inline unsigned blocks_for(int64_t n, int per = 256) { return (unsigned)((n + per - 1) / per); }

// Grid of a persistent kernel: every workgroup resident at once (compute
// units x workgroups per unit from the occupancy calculator), so that the
// strided tile loop never leaves a second, partial wave.
unsigned resident_grid(const void *kern, int threads, size_t lds, int64_t ntiles) {
    int dev = 0, cus = 0, per_cu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, lds) != hipSuccess || per_cu < 1)
        per_cu = 1;
    return (unsigned)std::min<int64_t>(ntiles, (int64_t)std::max(cus, 8) * per_cu);
}

// one pass of sub-FFTs of length L: persistent when the resident grid is
// smaller than the tile count (register prefetch of the next tile), else one
// tile per workgroup (no prefetch)
template <typename T, int L, int FORM, bool TAB>
void fft_form(const Geom &g, const FftPass &fp, unsigned nz, hipStream_t st, SigAddr<const cx<T>> in,
              SigAddr<cx<T>> out, const cx<T> *tw, const c64 *tlo, const c64 *thi,
              bool force_single) {
    constexpr int C = fft_cols<T, L>();
    // tiles of all nz signals (a batch is one persistent grid)
    const int64_t ntot = (fp.F * g.P + C - 1) / C * (int64_t)nz;
    const dim3 block(kFftThreads);
    // complex64 pass 1 over contiguous signals: two-column 16-byte loads
    // where every such load is 16-byte aligned (adjacent columns, even
    // strides, an aligned base)
    if constexpr (FORM == FFT_PASS1 && sizeof(T) == 4 && !TAB && C % 2 == 0) {
        const bool v2 = fp.in_fs == 1 && fp.in_es % 2 == 0 && fp.in_bs % 2 == 0 && g.M % 2 == 0 &&
                        reinterpret_cast<uintptr_t>(in.base) % 16 == 0 && (nz == 1 || in.stride % 2 == 0);
        if (v2 && !force_single) {
            const unsigned r2 =
                resident_grid((const void *)k_fft<T, L, FORM, C, true, TAB, true>, kFftThreads, 0, ntot);
            if ((int64_t)r2 < ntot)
                hipLaunchKernelGGL((k_fft<T, L, FORM, C, true, TAB, true>), dim3(r2, 1, 1), block, 0, st, g, fp, in,
                                   out, tw, tlo, thi, (int)nz);
            else
                hipLaunchKernelGGL((k_fft<T, L, FORM, C, false, TAB, true>), dim3((unsigned)ntot, 1, 1), block, 0, st,
                                   g, fp, in, out, tw, tlo, thi, (int)nz);
            return;
        }
    }
    // complex64 last pass into contiguous signals: two-column 16-byte stores
    // where every such store is 16-byte aligned
    if constexpr (FORM == FFT_PASS2 && sizeof(T) == 4 && !TAB && C % 2 == 0) {
        const bool s2 = fp.out_fs == 1 && fp.out_tile == 0 && fp.out_ks % 2 == 0 && !fp.twiddle && g.M % 2 == 0 &&
                        reinterpret_cast<uintptr_t>(out.base) % 16 == 0 && (nz == 1 || out.stride % 2 == 0);
        if (s2 && !force_single) {
            const unsigned r2 =
                resident_grid((const void *)k_fft<T, L, FORM, C, true, TAB, false, true>, kFftThreads, 0, ntot);
            if ((int64_t)r2 < ntot)
                hipLaunchKernelGGL((k_fft<T, L, FORM, C, true, TAB, false, true>), dim3(r2, 1, 1), block, 0, st, g, fp,
                                   in, out, tw, tlo, thi, (int)nz);
            else
                hipLaunchKernelGGL((k_fft<T, L, FORM, C, false, TAB, false, true>), dim3((unsigned)ntot, 1, 1), block, 0,
                                   st, g, fp, in, out, tw, tlo, thi, (int)nz);
            return;
        }
    }
    const unsigned res =
        force_single ? (unsigned)ntot : resident_grid((const void *)k_fft<T, L, FORM, C, true, TAB>, kFftThreads, 0, ntot);
    if ((int64_t)res < ntot)
        hipLaunchKernelGGL((k_fft<T, L, FORM, C, true, TAB>), dim3(res, 1, 1), block, 0, st, g, fp, in, out,
                           tw, tlo, thi, (int)nz);
    else
        hipLaunchKernelGGL((k_fft<T, L, FORM, C, false, TAB>), dim3((unsigned)ntot, 1, 1), block, 0, st, g, fp, in, out,
                           tw, tlo, thi, (int)nz);
}

template <typename T, int L, bool TAB>
void fft_len(const Geom &g, const FftPass &fp, unsigned nz, hipStream_t st, SigAddr<const cx<T>> in,
             SigAddr<cx<T>> out, const cx<T> *tw, const c64 *tlo, const c64 *thi) {
    if (fp.in_es == 1 && fp.out_ks == 1)
        fft_form<T, L, FFT_SINGLE, TAB>(g, fp, nz, st, in, out, tw, tlo, thi, false);
    else if (fp.in_es == 1)
        // FP64 pass 2 measures faster as one tile per workgroup (no
        // persistence): its strided 64-byte output pieces gain from many
        // workgroups in flight more than from the register prefetch
        fft_form<T, L, FFT_PASS2, TAB>(g, fp, nz, st, in, out, tw, tlo, thi, sizeof(T) == 8);
    else
        fft_form<T, L, FFT_PASS1, TAB>(g, fp, nz, st, in, out, tw, tlo, thi, false);
}

template <typename T, bool TAB>
void launch_fft(int64_t L, const Geom &g, const FftPass &fp, unsigned nz, hipStream_t st, SigAddr<const cx<T>> in,
                SigAddr<cx<T>> out, const cx<T> *tw, const c64 *tlo,
                const c64 *thi) {
    switch (L) {
        case 16: fft_len<T, 16, TAB>(g, fp, nz, st, in, out, tw, tlo, thi); break;
        case 32: fft_len<T, 32, TAB>(g, fp, nz, st, in, out, tw, tlo, thi); break;
        case 64: fft_len<T, 64, TAB>(g, fp, nz, st, in, out, tw, tlo, thi); break;
        case 128: fft_len<T, 128, TAB>(g, fp, nz, st, in, out, tw, tlo, thi); break;
        case 256: fft_len<T, 256, TAB>(g, fp, nz, st, in, out, tw, tlo, thi); break;
        case 512: fft_len<T, 512, TAB>(g, fp, nz, st, in, out, tw, tlo, thi); break;
        default: fft_len<T, 1024, TAB>(g, fp, nz, st, in, out, tw, tlo, thi); break;
    }
}

template <typename T, bool TAB>
void launch_combine(const Geom &g, unsigned nz, hipStream_t st, const cx<T> *psec, int64_t ps, const c64 *sumsec,
                    int64_t ss, const c64 *gfac, const c64 *wP, SigAddr<cx<T>> y) {
#define COMBINE(PP)                                                                                          \
    hipLaunchKernelGGL((k_combine<T, PP, TAB>),                                                                   \
                       dim3(resident_grid((const void *)k_combine<T, PP, TAB>, kCombineThreads, 0,                  \
                                          (g.M + combine_rows<PP>() - 1) / combine_rows<PP>()),                \
                            1, nz),                                                                          \
                       dim3(kCombineThreads), 0, st, g, psec, ps, sumsec, ss, gfac, wP, y)
    switch (g.P) {
        case 2: COMBINE(2); break;
        case 4: COMBINE(4); break;
        case 8: COMBINE(8); break;
        case 16: COMBINE(16); break;
        case 32: COMBINE(32); break;
        default: COMBINE(64); break;
    }
#undef COMBINE
}

template <typename T, int NRT, int G, int KT, int CT, bool TAB>
void leaf_mma_inst(const Geom &g, const Shapes &sh, unsigned nz, hipStream_t st, SigAddr<const cx<T>> x, const T *psi,
                   int64_t wstride_T, cx<T> *psec, int64_t wstride_cx, const TabPtrs<T> &tb, int conj) {
    auto kern = k_leaf_mma<T, NRT, G, KT, CT, TAB>;
    // persistent grid: as many workgroups as can be resident at once
    int dev = 0, cus = 0, per_cu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (sh.leaf_lds > 64 * 1024)
        (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh.leaf_lds);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, sh.leaf_lds) != hipSuccess || per_cu < 1)
        per_cu = 1;
    // persistent over the tiles of all nz signals; a multiple of the section
    // groups: each workgroup keeps one group
    const int64_t ngroups = g.P / G;
    const int64_t ntot = ngroups * (g.s / (8 * CT)) * (int64_t)nz;
    int64_t nwg64 = std::min<int64_t>(ntot, (int64_t)std::max(cus, 8) * per_cu);
    nwg64 = std::max<int64_t>(ngroups, nwg64 / ngroups * ngroups);
    const unsigned nwg = (unsigned)nwg64;
    hipLaunchKernelGGL(kern, dim3(nwg, 1, 1), dim3(256), sh.leaf_lds, st, g, x, psi, wstride_T, psec,
                       wstride_cx, tb.evfrag, tb.evalm, tb.evalmh, tb.kv, conj, sh.rm, (int)nz);
}

template <typename T, int NRT, int G, bool TAB>
void leaf_mma_kt(const Geom &g, const Shapes &sh, unsigned nz, hipStream_t st, SigAddr<const cx<T>> x, const T *psi,
                 int64_t wstride_T, cx<T> *psec, int64_t wstride_cx, const TabPtrs<T> &tb, int conj) {
    // column tiles: NRT 1 -> up to 4, NRT 2 -> up to 2 (leaf_ct); s >= 8
    const int ct = leaf_ct(g);
    if (leaf_kt(g.terms) == 4) {
        if (NRT == 1 && ct == 4)
            leaf_mma_inst<T, NRT, G, 4, (NRT == 1 ? 4 : 1), TAB>(g, sh, nz, st, x, psi, wstride_T, psec, wstride_cx, tb, conj);
        else if (NRT <= 2 && ct >= 2)
            leaf_mma_inst<T, NRT, G, 4, (NRT <= 2 ? 2 : 1), TAB>(g, sh, nz, st, x, psi, wstride_T, psec, wstride_cx, tb, conj);
        else
            leaf_mma_inst<T, NRT, G, 4, 1, TAB>(g, sh, nz, st, x, psi, wstride_T, psec, wstride_cx, tb, conj);
    } else {
        if (NRT == 1 && ct == 4)
            leaf_mma_inst<T, NRT, G, 8, (NRT == 1 ? 4 : 1), TAB>(g, sh, nz, st, x, psi, wstride_T, psec, wstride_cx, tb, conj);
        else if (NRT <= 2 && ct >= 2)
            leaf_mma_inst<T, NRT, G, 8, (NRT <= 2 ? 2 : 1), TAB>(g, sh, nz, st, x, psi, wstride_T, psec, wstride_cx, tb, conj);
        else
            leaf_mma_inst<T, NRT, G, 8, 1, TAB>(g, sh, nz, st, x, psi, wstride_T, psec, wstride_cx, tb, conj);
    }
}

template <typename T, int NRT, bool TAB>
void leaf_mma_g(const Geom &g, const Shapes &sh, unsigned nz, hipStream_t st, SigAddr<const cx<T>> x, const T *psi,
                int64_t wstride_T, cx<T> *psec, int64_t wstride_cx, const TabPtrs<T> &tb, int conj) {
    if (sh.leaf_G >= 4)
        leaf_mma_kt<T, NRT, 4, TAB>(g, sh, nz, st, x, psi, wstride_T, psec, wstride_cx, tb, conj);
    else
        leaf_mma_kt<T, NRT, 2, TAB>(g, sh, nz, st, x, psi, wstride_T, psec, wstride_cx, tb, conj);
}

template <typename T, bool TAB>
void launch_leaf_mma(const Geom &g, const Shapes &sh, unsigned nz, hipStream_t st, SigAddr<const cx<T>> x, const T *psi,
                     int64_t wstride_T, cx<T> *psec, int64_t wstride_cx, const TabPtrs<T> &tb, int conj) {
    switch (g.b / 16) {
        case 1: leaf_mma_g<T, 1, TAB>(g, sh, nz, st, x, psi, wstride_T, psec, wstride_cx, tb, conj); break;
        case 2: leaf_mma_g<T, 2, TAB>(g, sh, nz, st, x, psi, wstride_T, psec, wstride_cx, tb, conj); break;
        case 4: leaf_mma_g<T, 4, TAB>(g, sh, nz, st, x, psi, wstride_T, psec, wstride_cx, tb, conj); break;
        default: leaf_mma_g<T, 8, TAB>(g, sh, nz, st, x, psi, wstride_T, psec, wstride_cx, tb, conj); break;
    }
}

template <int NRT, int G, int KT, bool TAB>
void leaf_h16_inst(const fmmfft_plan_s *p, unsigned nz, hipStream_t st, SigAddr<const c32> x, const float *psi,
                   int64_t wstride_T, c32 *psec, int64_t wstride_cx, const TabPtrs<float> &tb, int conj) {
    const Geom &g = p->prm.g;
    const Shapes &sh = p->sh;
    auto kern = k_leaf_h16<NRT, G, KT, TAB>;
    int dev = 0, cus = 0, per_cu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (sh.leaf_lds > 64 * 1024)
        (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh.leaf_lds);
    constexpr int nth = 64 * (G > 4 ? G : 4);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, nth, sh.leaf_lds) != hipSuccess || per_cu < 1)
        per_cu = 1;
    // persistent over the tiles of all nz signals; a multiple of the section
    // groups: each workgroup keeps one group
    const int64_t ngroups = g.P / G;
    const int64_t ntot = ngroups * (g.s / LeafH16<NRT>::BB) * (int64_t)nz;
    int64_t nwg64 = std::min<int64_t>(ntot, (int64_t)std::max(cus, 8) * per_cu);
    nwg64 = std::max<int64_t>(ngroups, nwg64 / ngroups * ngroups);
    hipLaunchKernelGGL(kern, dim3((unsigned)nwg64, 1, 1), dim3(nth), sh.leaf_lds, st, g, x, psi, wstride_T,
                       psec, wstride_cx, tb.evfrag, tb.evalm, tb.evalmh, tb.kv,
                       reinterpret_cast<const h16x4 *>(p->kfrag), p->tab64 + p->tl.kv, conj, sh.rm, (int)nz);
}

template <int NRT, int G, bool TAB>
void leaf_h16_kt(const fmmfft_plan_s *p, unsigned nz, hipStream_t st, SigAddr<const c32> x, const float *psi,
                 int64_t wstride_T, c32 *psec, int64_t wstride_cx, const TabPtrs<float> &tb, int conj) {
    if (leaf_kt(p->prm.g.terms) == 4)
        leaf_h16_inst<NRT, G, 4, TAB>(p, nz, st, x, psi, wstride_T, psec, wstride_cx, tb, conj);
    else
        leaf_h16_inst<NRT, G, 8, TAB>(p, nz, st, x, psi, wstride_T, psec, wstride_cx, tb, conj);
}

template <bool TAB>
void launch_leaf_h16(const fmmfft_plan_s *p, unsigned nz, hipStream_t st, SigAddr<const c32> x, const float *psi,
                     int64_t wstride_T, c32 *psec, int64_t wstride_cx, const TabPtrs<float> &tb, int conj) {
    const int G = p->sh.leaf_G;
    if (p->prm.g.b == 64) {
        if (G >= 8) leaf_h16_kt<4, 8, TAB>(p, nz, st, x, psi, wstride_T, psec, wstride_cx, tb, conj);
        else if (G >= 4) leaf_h16_kt<4, 4, TAB>(p, nz, st, x, psi, wstride_T, psec, wstride_cx, tb, conj);
        else leaf_h16_kt<4, 2, TAB>(p, nz, st, x, psi, wstride_T, psec, wstride_cx, tb, conj);
    } else {
        if (G >= 8) leaf_h16_kt<2, 8, TAB>(p, nz, st, x, psi, wstride_T, psec, wstride_cx, tb, conj);
        else if (G >= 4) leaf_h16_kt<2, 4, TAB>(p, nz, st, x, psi, wstride_T, psec, wstride_cx, tb, conj);
        else leaf_h16_kt<2, 2, TAB>(p, nz, st, x, psi, wstride_T, psec, wstride_cx, tb, conj);
    }
}

// The kernel's discarded stores (section 0, padding terms) go to sumsec,
// which the section-sum reduction launched after it overwrites.
template <typename T, int VEC, int MT, bool TAB>
void s2m_mma_ks(const Geom &g, const Shapes &sh, unsigned nz, hipStream_t st, SigAddr<const cx<T>> x, T *phi,
                int64_t wstride_T, c64 *sums_part, int64_t wstride_c64, const T *initch, const T *opfrag,
                c64 *sumsec, int conj) {
    const dim3 grid(sh.s2m_nblk, 1, nz);
    const bool quad = sh.s2m_levels == 2;
    T *sink = reinterpret_cast<T *>(sumsec);
    const int64_t sink_stride = wstride_c64 * 16 / (int64_t)sizeof(T);
    switch (g.b) {
        case 16:
            if (quad)
                hipLaunchKernelGGL((k_s2m_mma<T, VEC, MT, 4, true, TAB>), grid, dim3(256), 0, st, g, x, phi,
                                   wstride_T, sums_part, wstride_c64, initch, opfrag, sink, sink_stride, conj);
            else
                hipLaunchKernelGGL((k_s2m_mma<T, VEC, MT, 4, false, TAB>), grid, dim3(256), 0, st, g, x, phi,
                                   wstride_T, sums_part, wstride_c64, initch, opfrag, sink, sink_stride, conj);
            break;
        case 32:
            if (quad)
                hipLaunchKernelGGL((k_s2m_mma<T, VEC, MT, 8, true, TAB>), grid, dim3(256), 0, st, g, x, phi,
                                   wstride_T, sums_part, wstride_c64, initch, opfrag, sink, sink_stride, conj);
            else
                hipLaunchKernelGGL((k_s2m_mma<T, VEC, MT, 8, false, TAB>), grid, dim3(256), 0, st, g, x, phi,
                                   wstride_T, sums_part, wstride_c64, initch, opfrag, sink, sink_stride, conj);
            break;
        default:
            if (quad)
                hipLaunchKernelGGL((k_s2m_mma<T, VEC, MT, 16, true, TAB>), grid, dim3(256), 0, st, g, x, phi,
                                   wstride_T, sums_part, wstride_c64, initch, opfrag, sink, sink_stride, conj);
            else
                hipLaunchKernelGGL((k_s2m_mma<T, VEC, MT, 16, false, TAB>), grid, dim3(256), 0, st, g, x, phi,
                                   wstride_T, sums_part, wstride_c64, initch, opfrag, sink, sink_stride, conj);
            break;
    }
}

// matrix-core S2M (complex64): VEC from the plan's shapes, MT = 16-term blocks
template <typename T, bool TAB>
void launch_s2m_mma(const Geom &g, const Shapes &sh, unsigned nz, hipStream_t st, SigAddr<const cx<T>> x, T *phi,
                    int64_t wstride_T, c64 *sums_part, int64_t wstride_c64, const T *initch, const T *opfrag,
                    c64 *sumsec, int conj) {
    if constexpr (sizeof(T) == 4) {
        if (sh.s2m_vec == 4) {
            if (g.terms <= 16) s2m_mma_ks<T, 4, 1, TAB>(g, sh, nz, st, x, phi, wstride_T, sums_part, wstride_c64, initch, opfrag, sumsec, conj);
            else s2m_mma_ks<T, 4, 2, TAB>(g, sh, nz, st, x, phi, wstride_T, sums_part, wstride_c64, initch, opfrag, sumsec, conj);
        } else {
            if (g.terms <= 16) s2m_mma_ks<T, 2, 1, TAB>(g, sh, nz, st, x, phi, wstride_T, sums_part, wstride_c64, initch, opfrag, sumsec, conj);
            else s2m_mma_ks<T, 2, 2, TAB>(g, sh, nz, st, x, phi, wstride_T, sums_part, wstride_c64, initch, opfrag, sumsec, conj);
        }
    }
}

template <typename T, int PT, bool TAB, typename Mark>
void s2m_tree_pt(const Geom &g, const Shapes &sh, unsigned nz, hipStream_t st, SigAddr<const cx<T>> x, T *phi, T *psi,
                 int64_t wstride_T, c64 *sums_part, int64_t wstride_c64, c64 *sumsec, const TabPtrs<T> &tb,
                 int conj, Mark &mark) {
    if (sh.s2m_vec)
        launch_s2m_mma<T, TAB>(g, sh, nz, st, x, phi, wstride_T, sums_part, wstride_c64, tb.initch, tb.opfrag, sumsec,
                          conj);
    else
        hipLaunchKernelGGL((k_s2m<T, PT, TAB>), dim3(sh.s2m_nblk, 1, nz), dim3(kS2mThreads), sh.s2m_lds, st, g, x,
                           phi, wstride_T, sums_part, wstride_c64, tb.initch, conj);
    // tree passes on the matrix cores: terms padded to 16 or 32, one wave per
    // (node, block of 16 rows), four waves per workgroup
    constexpr int TP = PT <= 16 ? 16 : 32;
    const int64_t nrb = (tree_rows(g) + 15) / 16;
    // nodes of at most two row blocks (P <= 16): packed form, several
    // nodes per wave
    const bool pack = nrb <= 2;
    // the matrix-core S2M has already formed levels n-1 (and n-2)
    int lu = g.n - 1 - (sh.s2m_vec ? sh.s2m_levels : 0);
    // complex64, 16 terms, nodes of more than two row blocks: upward levels
    // two per launch, the first of which also reduces the section sums
    // (also for 32-row nodes, P = 16: 256 x 2^20 batch 6.30 -> 6.25 ms)
    const bool up2 = sizeof(T) == 4 && TP == 16 && (!pack || tree_rows(g) == 32) && lu - 1 >= 2;
    const SumsJob sj{sums_part, (int)sh.s2m_nblk, wstride_c64, sumsec, wstride_c64};
    // downward levels: the split-product M2L instances for FMMFFT_C64_SPLIT
    const bool split = tb.opfrag_lo != nullptr;
    if (!up2) hipLaunchKernelGGL(k_sums_reduce, dim3(g.P, 1, nz), dim3(256), 0, st, g, sj);
    mark(1);
    auto grid = [&](int l, int mode) {
        const int64_t nn = int64_t(1) << l;
        int64_t waves;
        if (pack) {
            const int64_t npw = tree_mma_rb_packed<T, TP>(mode) / nrb;
            waves = 2 * ((nn + 2 * npw - 1) / (2 * npw));
        } else {
            const int rb = tree_mma_rb<T, TP>(mode);
            waves = nn * ((nrb + rb - 1) / rb);
        }
        return dim3((unsigned)((waves + 3) / 4), 1, nz);
    };
    auto level = [&](auto kern, auto pkern, int l, int mode) {
        hipLaunchKernelGGL(pack ? pkern : kern, grid(l, mode), dim3(256), 0, st, g, l, (const T *)phi, phi,
                           (const T *)psi, psi, wstride_T, tb.opfrag, tb.opfrag_lo);
    };
    // complex64 16-term trees whose nodes are whole 64-row groups (P >= 32):
    // the finest downward levels in the sibling form
    const bool sib4 = sizeof(T) == 4 && TP == 16 && !pack && tree_rows(g) % 64 == 0;
    auto down = [&](int l, int mode) {
        if (mode == 2) {
            if (split) level(k_tree_mma<T, TP, 2, false, true>, k_tree_mma<T, TP, 2, true, true>, l, 2);
            else level(k_tree_mma<T, TP, 2, false>, k_tree_mma<T, TP, 2, true>, l, 2);
        } else {
            if (split) level(k_tree_mma<T, TP, 1, false, true>, k_tree_mma<T, TP, 1, true, true>, l, 1);
            else level(k_tree_mma<T, TP, 1, false>, k_tree_mma<T, TP, 1, true>, l, 1);
        }
    };
    // levels (l, l - 1) two per launch; the launch that ends at level 2 also
    // applies the top flip (not in the split-M2L form, whose top flip takes
    // the split operator)
    bool top_done = false;
    if constexpr (sizeof(T) == 4 && TP == 16) {
        if (up2)
            for (bool first = true; lu - 1 >= 2; lu -= 2, first = false) {
                const unsigned nt = (unsigned)(((int64_t(1) << (lu - 1)) * nrb + 3) / 4);
                const bool topf = lu - 1 == 2 && !split;
                const dim3 gd(nt + (first ? (unsigned)g.P : 0u), 1, nz);
                if (first && topf)
                    hipLaunchKernelGGL((k_tree_up2<T, TP, true, true>), gd, dim3(256), 0, st, g, lu, phi, wstride_T,
                                       tb.opfrag, psi, sj);
                else if (first)
                    hipLaunchKernelGGL((k_tree_up2<T, TP, true, false>), gd, dim3(256), 0, st, g, lu, phi, wstride_T,
                                       tb.opfrag, psi, sj);
                else if (topf)
                    hipLaunchKernelGGL((k_tree_up2<T, TP, false, true>), gd, dim3(256), 0, st, g, lu, phi, wstride_T,
                                       tb.opfrag, psi, sj);
                else
                    hipLaunchKernelGGL((k_tree_up2<T, TP, false, false>), gd, dim3(256), 0, st, g, lu, phi,
                                       wstride_T, tb.opfrag, psi, sj);
                top_done = topf;
            }
    }
    for (; lu >= 2; --lu) level(k_tree_mma<T, TP, 0, false>, k_tree_mma<T, TP, 0, true>, lu, 0);
    mark(2);
    if (!top_done) down(2, 2);
    int l = 3;
    if constexpr (sizeof(T) == 4 && TP == 16) {
        // complex64, 16 terms, nodes of more than two row blocks: levels
        // 3..last in pairs (l, l + 1) (an odd count leaves level 3 on its
        // own), the finer ones one per launch: the two finest as a pair
        // measured slower (0.181 vs 0.163 ms at N = 2^26), and with whole
        // 64-row groups the sibling form (k_tree_mma SIB4) takes every level
        // from 11 on (levels 11..14 at N = 2^26: 0.176 ms downward against
        // 0.186 with pairs up to level 12)
        if (!pack) {
            const int last = sib4 ? std::min(g.n - 2, 10) : g.n - 2;
            if ((last - 2) % 2 == 1) down(l++, 1);
            for (; l + 1 <= last; l += 2) {
                const dim3 gd((unsigned)(((int64_t(1) << l) * nrb + 3) / 4), 1, nz);
                if (split)
                    hipLaunchKernelGGL((k_tree_down2<T, TP, true>), gd, dim3(256), 0, st, g, l, (const T *)phi, psi,
                                       wstride_T, tb.opfrag, tb.opfrag_lo);
                else
                    hipLaunchKernelGGL((k_tree_down2<T, TP, false>), gd, dim3(256), 0, st, g, l, (const T *)phi, psi,
                                       wstride_T, tb.opfrag, tb.opfrag_lo);
            }
        }
    }
    // complex64 16-term trees of 32-row nodes (P = 16): every downward level
    // from 3 on in the sibling form with two rows per lane (five source
    // nodes per pair instead of the packed form's eight: 256 x 2^20, P = 16,
    // downward 1.21 -> 1.09 ms)
    const bool sib2 = sizeof(T) == 4 && TP == 16 && pack && tree_rows(g) == 32;
    for (; l <= g.n; ++l) {
        if constexpr (sizeof(T) == 4 && TP == 16) {
            if (sib2) {
                const dim3 gd((unsigned)(((int64_t(1) << (l - 1)) + 3) / 4), 1, nz);
                if (split)
                    hipLaunchKernelGGL((k_tree_mma<T, TP, 1, false, true, true, 2>), gd, dim3(256), 0, st, g, l,
                                       (const T *)phi, phi, (const T *)psi, psi, wstride_T, tb.opfrag, tb.opfrag_lo);
                else
                    hipLaunchKernelGGL((k_tree_mma<T, TP, 1, false, false, true, 2>), gd, dim3(256), 0, st, g, l,
                                       (const T *)phi, phi, (const T *)psi, psi, wstride_T, tb.opfrag, tb.opfrag_lo);
                continue;
            }
            if (sib4) {
                // one wave per sibling pair and 64-row group
                const dim3 gd((unsigned)(((int64_t(1) << (l - 1)) * (tree_rows(g) / 64) + 3) / 4), 1, nz);
                if (split)
                    hipLaunchKernelGGL((k_tree_mma<T, TP, 1, false, true, true>), gd, dim3(256), 0, st, g, l,
                                       (const T *)phi, phi, (const T *)psi, psi, wstride_T, tb.opfrag,
                                       tb.opfrag_lo);
                else
                    hipLaunchKernelGGL((k_tree_mma<T, TP, 1, false, false, true>), gd, dim3(256), 0, st, g, l,
                                       (const T *)phi, phi, (const T *)psi, psi, wstride_T, tb.opfrag,
                                       tb.opfrag_lo);
                continue;
            }
        }
        down(l, 1);
    }
}

template <typename T, bool TAB, typename Mark>
void launch_s2m_tree(int pt, const fmmfft_plan_s *, const Geom &g, const Shapes &sh, unsigned nz, hipStream_t st,
                     SigAddr<const cx<T>> x, T *phi, T *psi, int64_t wstride_T, c64 *sums_part, int64_t wstride_c64,
                     c64 *sumsec, const TabPtrs<T> &tb, int conj, Mark &mark) {
    switch (pt) {
        case 8: s2m_tree_pt<T, 8, TAB>(g, sh, nz, st, x, phi, psi, wstride_T, sums_part, wstride_c64, sumsec, tb, conj, mark); break;
        case 16: s2m_tree_pt<T, 16, TAB>(g, sh, nz, st, x, phi, psi, wstride_T, sums_part, wstride_c64, sumsec, tb, conj, mark); break;
        case 24: s2m_tree_pt<T, 24, TAB>(g, sh, nz, st, x, phi, psi, wstride_T, sums_part, wstride_c64, sumsec, tb, conj, mark); break;
        default: s2m_tree_pt<T, 32, TAB>(g, sh, nz, st, x, phi, psi, wstride_T, sums_part, wstride_c64, sumsec, tb, conj, mark); break;
    }
}

// One execute of count signals: their inputs and outputs at x.at(sig) /
// out.at(sig) (contiguous, or a device table of per-signal addresses),
// workspace ws (count strides).
template <typename T, bool TAB>
int run(const fmmfft_plan_s *p, int64_t count, SigAddr<const cx<T>> x, SigAddr<cx<T>> out, int dir, char *ws,
        hipStream_t st) {
    const Geom &g = p->prm.g;
    const Shapes &sh = p->sh;
    const WsLayout &wl = p->ws;
    const TabPtrs<T> tb = tab_ptrs<T>(p);
    const int conj = dir < 0 ? 1 : 0;
    const unsigned nz = (unsigned)count;
    const int64_t wstride_T = (int64_t)(wl.stride / sizeof(T));
    const int64_t wstride_c64 = (int64_t)(wl.stride / sizeof(c64));
    const int64_t wstride_cx = (int64_t)(wl.stride / sizeof(cx<T>));

    // the same output buffers as FFT input, and the workspace rows as
    // signal addresses
    const SigAddr<const cx<T>> out_in{out.base, out.stride, out.tab};
    T *phi = reinterpret_cast<T *>(ws + wl.phi);
    T *psi = reinterpret_cast<T *>(ws + wl.psi);
    c64 *sums_part = reinterpret_cast<c64 *>(ws + wl.sums_part);
    c64 *sumsec = reinterpret_cast<c64 *>(ws + wl.sumsec);
    cx<T> *psec = reinterpret_cast<cx<T> *>(ws + wl.psec);
    const SigAddr<cx<T>> psec_a{psec, wstride_cx, nullptr};
    const SigAddr<const cx<T>> psec_in{psec, wstride_cx, nullptr};
    const c64 *gfac = reinterpret_cast<const c64 *>(p->tab64 + p->tl.gfac);
    const c64 *wP = reinterpret_cast<const c64 *>(p->tab64 + p->tl.wP);
    const c64 *tlo = reinterpret_cast<const c64 *>(p->tab64 + p->tl.tlo);
    const c64 *thi = reinterpret_cast<const c64 *>(p->tab64 + p->tl.thi);

    hipEvent_t *ev = nullptr;
    if (p->prof_ev && p->prof_calls < p->prof_cap) {
        ev = p->prof_ev + (size_t)p->prof_calls * (FMMFFT_NSTAGES + 1);
        const_cast<fmmfft_plan_s *>(p)->prof_calls++;
    }
    auto mark = [&](int i) {
        if (ev) (void)hipEventRecord(ev[i], st);
    };
    mark(0);
    launch_s2m_tree<T, TAB>(padded_terms(g.terms), p, g, sh, nz, st, x, phi, psi, wstride_T, sums_part, wstride_c64,
                       sumsec, tb, conj, mark);
    mark(3);
    bool leaf_done = false;
    if constexpr (sizeof(T) == 4) {
        if (sh.leaf_h16) {
            launch_leaf_h16<TAB>(p, nz, st, x, psi, wstride_T, psec, wstride_cx, tb, conj);
            leaf_done = true;
        }
    }
    if (leaf_done) {
    } else if (sh.leaf_mma) {
        launch_leaf_mma<T, TAB>(g, sh, nz, st, x, psi, wstride_T, psec, wstride_cx, tb, conj);
    } else {
        hipLaunchKernelGGL((k_leaf<T, TAB>), dim3(g.s / sh.leaf_bb, g.P / sh.leaf_G, nz), dim3(256), sh.leaf_lds, st, g,
                           x, psi, wstride_T, psec, wstride_cx, tb.evalm, tb.evalmh, tb.kv, sh.leaf_G,
                           sh.leaf_bb, conj, sh.rm);
    }
    mark(4);
    launch_combine<T, TAB>(g, nz, st, psec, wstride_cx, sumsec, wstride_c64, gfac, wP, out);
    mark(5);

    const FftLayout &f = p->fft;
    const cx<T> *twA = reinterpret_cast<const cx<T> *>(tb.twA);
    const cx<T> *twB = reinterpret_cast<const cx<T> *>(tb.twB);
    if (!f.two_pass) {
        FftPass fp{};
        fp.F = 1;
        fp.in_fs = g.M; fp.in_es = 1; fp.in_bs = 0; fp.in_ltw = 30;
        fp.out_fs = g.M; fp.out_ks = 1;
        fp.conj_out = conj;
        fp.logR = f.logR;
        launch_fft<T, TAB>(f.M1, g, fp, nz, st, out_in, out, twA, tlo, thi);
        mark(6);
    } else {
        // first pass: sub-FFTs of length M1 along the columns of [M1][R]
        const int64_t R = g.M / f.M1;
        FftPass a{};
        a.F = R;
        if (sh.rm.ltw >= 0) {
            // column-tiled input (the leaf's row map, tiles of TW >= C
            // columns): column f0 + j of tile q = f0 / TW at
            // q TW M1 + f0 % TW + j, element e at e TW
            a.in_fs = 1; a.in_es = int64_t(1) << sh.rm.ltw;
            a.in_bs = f.M1 << sh.rm.ltw; a.in_ltw = sh.rm.ltw;
        } else {
            a.in_fs = 1; a.in_es = R; a.in_bs = 0; a.in_ltw = 30;
        }
        a.out_fs = 1; a.out_ks = R;
        a.twiddle = 1;
        a.logR = f.logR;
        launch_fft<T, TAB>(f.M1, g, a, nz, st, out_in, psec_a, twA, tlo, thi);
        mark(6);
        if (f.passes == 2) {
            FftPass bq{};
            bq.F = f.M1;
            bq.in_fs = f.M2; bq.in_es = 1; bq.in_bs = 0; bq.in_ltw = 30;
            bq.out_fs = 1; bq.out_ks = f.M1;
            bq.conj_out = conj;
            bq.logR = f.logR;
            launch_fft<T, TAB>(f.M2, g, bq, nz, st, psec_in, out, twB, tlo, thi);
        } else {
            // Z[k1][r], r = n2 M3 + n3.  Second pass, in place: for every
            // (k1, n3) (sub-FFT f = k1 M3 + n3, a column of row k1) the
            // M2-point FFT over n2, times w_R^{n3 k2} = w_M^{M1 n3 k2}
            const int l1 = ilog2(f.M1), l3 = ilog2(f.M3);
            const cx<T> *twC = reinterpret_cast<const cx<T> *>(tb.twC);
            FftPass b2{};
            b2.F = f.M1 * f.M3;
            b2.in_fs = 1; b2.in_es = f.M3; b2.in_bs = R; b2.in_ltw = l3;
            b2.out_fs = 1; b2.out_ks = f.M3; b2.out_bs = R; b2.out_tile = l3;
            b2.twiddle = 1; b2.tw_lf = l3; b2.tw_shift = l1;
            b2.logR = f.logR;
            launch_fft<T, TAB>(f.M2, g, b2, nz, st, psec_in, psec_a, twB, tlo, thi);
            // third pass: for every (k1, k2) (sub-FFT f = k2 M1 + k1: k1
            // fastest, so that the outputs X[k1 + M1 k2 + M1 M2 k3] of
            // neighbouring sub-FFTs are neighbours) the M3-point FFT of the
            // contiguous row at k1 R + k2 M3
            FftPass b3{};
            b3.F = f.M1 * f.M2;
            b3.in_fs = f.M2 * f.M3; b3.in_es = 1; b3.in_bs = f.M3; b3.in_ltw = l1;
            b3.out_fs = 1; b3.out_ks = f.M1 * f.M2;
            b3.conj_out = conj;
            b3.logR = f.logR;
            launch_fft<T, TAB>(f.M3, g, b3, nz, st, psec_in, out, twC, tlo, thi);
        }
    }
    // complex64 with 12 terms or more: the P block DC bins from the FP64
    // section sums.  Those bins are then the exact DFT's, where the FMM-FFT
    // carries its truncation error (below 1.1e-11 of the result from 12
    // terms on, far under the complex64 rounding; with fewer terms the
    // transform keeps the FMM-FFT's own values)
    if constexpr (sizeof(T) == 4)
        if (g.terms >= 12)
            hipLaunchKernelGGL((k_dc_exact<T, TAB>), dim3(1, 1, nz), dim3(64), 0, st, g, (const c64 *)sumsec,
                               wstride_c64, wP, out, conj);
    mark(7);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        g_last_hip_error.store((int)e);
        return FMMFFT_EHIP;
    }
    return FMMFFT_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

#ifndef FMMFFT_TAB_UNIT
extern "C" {

int fmmfft_plan_sizes(int64_t n, int32_t p_blocks, int32_t terms, int32_t leaf, int32_t dtype,
                      size_t *table_bytes, size_t *work_bytes) {
    Params prm{};
    int rc = validate(n, p_blocks, terms, leaf, dtype, &prm);
    if (rc != FMMFFT_OK) return rc;
    const FftLayout f = fft_layout(prm.g);
    const TableLayout tl = table_layout(prm.g, f);
    const Shapes sh = shapes(prm.g, prm.esize);
    const WsLayout wl = ws_layout(prm.g, prm.esize, sh);
    if (table_bytes) {
        size_t t = tl.total * sizeof(double);
        const size_t tp = 4 * (size_t)leaf_kt(prm.g.terms), nopf = ((size_t)(prm.g.n - 2) * NOPS + 1) * tp * tp;
        if (is_c64(dtype)) t += (tl.n_ops + 2 * f.M1 + 2 * f.M2 + 2 * f.M3) * sizeof(float);
        if (dtype == FMMFFT_C64_SPLIT) t += nopf * sizeof(float);
        if (is_c64(dtype) && sh.leaf_h16)
            t += (size_t)prm.g.S * (3 * prm.g.b / 16 + prm.g.b / 16 - 1) * 2 * 64 * 4 * sizeof(_Float16);
        *table_bytes = t;
    }
    if (work_bytes) *work_bytes = wl.stride;
    return FMMFFT_OK;
}

int fmmfft_plan_create(fmmfft_plan *out, int64_t n, int32_t p_blocks, int32_t terms, int32_t leaf,
                       int32_t dtype, int32_t max_batch) {
    return fmmfft::plan_create_impl(out, n, p_blocks, terms, leaf, dtype, max_batch, true);
}

}  // extern "C"

// Plan creation; with_workspace = false skips the default workspace (for
// callers that always pass their own, fmmfft_execute_ws).
int fmmfft::plan_create_impl(fmmfft_plan *out, int64_t n, int32_t p_blocks, int32_t terms, int32_t leaf,
                             int32_t dtype, int32_t max_batch, bool with_workspace) {
    if (!out) return FMMFFT_EINVAL;
    *out = nullptr;
    if (max_batch < 0) return FMMFFT_EINVAL;
    Params prm{};
    int rc = validate(n, p_blocks, terms, leaf, dtype, &prm);
    if (rc != FMMFFT_OK) return rc;
    fmmfft_plan_s *p = new (std::nothrow) fmmfft_plan_s();
    if (!p) return FMMFFT_ENOMEM;
    p->prm = prm;
    p->fft = fft_layout(prm.g);
    p->tl = table_layout(prm.g, p->fft);
    p->sh = shapes(prm.g, prm.esize);
    p->ws = ws_layout(prm.g, prm.esize, p->sh);
    p->max_batch = max_batch == 0 ? 1 : max_batch;
    const Geom &g = prm.g;

    auto fail = [&](int code) {
        fmmfft_plan_destroy(p);
        return code;
    };
#define PLAN_TRY(expr)                                                                 \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                        \
            g_last_hip_error.store((int)e_);                                           \
            return fail(e_ == hipErrorOutOfMemory ? FMMFFT_ENOMEM : FMMFFT_EHIP);      \
        }                                                                              \
    } while (0)

    PLAN_TRY(hipGetDevice(&p->device));
    PLAN_TRY(hipMalloc(&p->tab64, p->tl.total * sizeof(double)));
    hipStream_t st;
    PLAN_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    double *t = p->tab64;
    const TableLayout &tl = p->tl;
    const int64_t n_closed = g.terms + (int64_t)g.b * g.terms + (int64_t)g.S * 4 * g.b + g.P;
    hipLaunchKernelGGL(k_build_closed, dim3(blocks_for(n_closed)), dim3(256), 0, st, g, t + tl.cheb,
                       t + tl.initch, t + tl.kv, reinterpret_cast<double2 *>(t + tl.gfac),
                       reinterpret_cast<double2 *>(t + tl.wP));
    const int jobs = 8 * (g.n - 2) + 1 + g.S * g.b + g.P / 2;
    hipLaunchKernelGGL(k_build_interp, dim3(jobs), dim3(64), 0, st, g, t + tl.ops, t + tl.topflp,
                       t + tl.evalm, t + tl.evalmh);
    if (g.b % 16 == 0) {
        const int nrt = g.b / 16, kt = leaf_kt(g.terms);
        hipLaunchKernelGGL(k_build_evfrag, dim3(blocks_for((int64_t)g.S * nrt * kt * 64)), dim3(256), 0, st, g,
                           (const double *)(t + tl.evalm), t + tl.evfrag, nrt, kt);
    }
    {
        const int tp = 4 * leaf_kt(g.terms);
        const int64_t nf = ((int64_t)(g.n - 2) * NOPS + 1) * tp * tp;
        hipLaunchKernelGGL(k_build_opfrag, dim3(blocks_for(nf)), dim3(256), 0, st, g, (const double *)(t + tl.ops),
                           (const double *)(t + tl.topflp), t + tl.opfrag, tp);
    }
    const int64_t R = int64_t(1) << p->fft.logR;
    const int64_t MA = p->fft.M1, MB = p->fft.M2, MC = p->fft.M3;
    hipLaunchKernelGGL(k_build_twiddles, dim3(blocks_for(MA + R + g.M / R)), dim3(256), 0, st, MA, g.M, R,
                       reinterpret_cast<double2 *>(t + tl.twA), reinterpret_cast<double2 *>(t + tl.tlo),
                       reinterpret_cast<double2 *>(t + tl.thi));
    if (p->fft.two_pass)
        hipLaunchKernelGGL(k_build_twiddles, dim3(blocks_for(MB)), dim3(256), 0, st, MB, g.M, int64_t(0),
                           reinterpret_cast<double2 *>(t + tl.twB), reinterpret_cast<double2 *>(t + tl.tlo),
                           reinterpret_cast<double2 *>(t + tl.thi));
    if (p->fft.passes == 3)
        hipLaunchKernelGGL(k_build_twiddles, dim3(blocks_for(MC)), dim3(256), 0, st, MC, g.M, int64_t(0),
                           reinterpret_cast<double2 *>(t + tl.twC), reinterpret_cast<double2 *>(t + tl.tlo),
                           reinterpret_cast<double2 *>(t + tl.thi));
    if (is_c64(prm.dtype)) {
        const size_t nA = 2 * (size_t)MA, nB = 2 * (size_t)MB, nC = 2 * (size_t)MC;
        const size_t tp = 4 * (size_t)leaf_kt(g.terms);
        const size_t nopf = prm.dtype == FMMFFT_C64_SPLIT ? ((size_t)(g.n - 2) * NOPS + 1) * tp * tp : 0;
        p->tw32_A = tl.n_ops;
        p->tw32_B = tl.n_ops + nA;
        p->tw32_C = tl.n_ops + nA + nB;
        p->opf_lo32 = tl.n_ops + nA + nB + nC;
        const size_t n32 = tl.n_ops + nA + nB + nC + nopf;
        hipError_t e = hipMalloc(&p->tab32, n32 * sizeof(float));
        if (e != hipSuccess) {
            (void)hipStreamDestroy(st);
            g_last_hip_error.store((int)e);
            return fail(e == hipErrorOutOfMemory ? FMMFFT_ENOMEM : FMMFFT_EHIP);
        }
        hipLaunchKernelGGL(k_to_f32, dim3(blocks_for(tl.n_ops)), dim3(256), 0, st, t, p->tab32, (int64_t)tl.n_ops);
        hipLaunchKernelGGL(k_to_f32, dim3(blocks_for(nA)), dim3(256), 0, st, t + tl.twA, p->tab32 + p->tw32_A,
                           (int64_t)nA);
        if (p->fft.two_pass)
            hipLaunchKernelGGL(k_to_f32, dim3(blocks_for(nB)), dim3(256), 0, st, t + tl.twB,
                               p->tab32 + p->tw32_B, (int64_t)nB);
        if (p->fft.passes == 3)
            hipLaunchKernelGGL(k_to_f32, dim3(blocks_for(nC)), dim3(256), 0, st, t + tl.twC,
                               p->tab32 + p->tw32_C, (int64_t)nC);
        if (nopf)
            hipLaunchKernelGGL(k_to_f32_lo, dim3(blocks_for(nopf)), dim3(256), 0, st, t + tl.opfrag,
                               p->tab32 + p->opf_lo32, (int64_t)nopf);
    }
    if (is_c64(prm.dtype) && p->sh.leaf_h16) {
        const int nrt = g.b / 16, nf = 3 * g.b / 16 + nrt - 1;
        const int64_t nk = (int64_t)g.S * nf * 2 * 64 * 4;
        hipError_t e = hipMalloc(&p->kfrag, nk * sizeof(_Float16));
        if (e != hipSuccess) {
            (void)hipStreamDestroy(st);
            g_last_hip_error.store((int)e);
            return fail(e == hipErrorOutOfMemory ? FMMFFT_ENOMEM : FMMFFT_EHIP);
        }
        hipLaunchKernelGGL(k_build_kfrag, dim3(blocks_for(nk)), dim3(256), 0, st, g, (const double *)(t + tl.kv),
                           p->kfrag, nrt, nf);
    }
    hipError_t e1 = hipGetLastError();
    hipError_t e2 = hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    PLAN_TRY(e1);
    PLAN_TRY(e2);
    if (with_workspace) PLAN_TRY(hipMalloc(&p->work, p->ws.stride * (size_t)p->max_batch));
#undef PLAN_TRY
    *out = p;
    return FMMFFT_OK;
}

extern "C" {

static void prof_free(fmmfft_plan plan) {
    if (plan->prof_ev) {
        for (size_t i = 0; i < (size_t)plan->prof_cap * (FMMFFT_NSTAGES + 1); ++i)
            (void)hipEventDestroy(plan->prof_ev[i]);
        delete[] plan->prof_ev;
    }
    plan->prof_ev = nullptr;
    plan->prof_cap = 0;
    plan->prof_calls = 0;
}

int fmmfft_profile_begin(fmmfft_plan plan, int32_t max_calls) {
    if (!plan || max_calls <= 0 || max_calls > 100000) return FMMFFT_EINVAL;
    prof_free(plan);
    const size_t n = (size_t)max_calls * (FMMFFT_NSTAGES + 1);
    plan->prof_ev = new (std::nothrow) hipEvent_t[n];
    if (!plan->prof_ev) return FMMFFT_ENOMEM;
    for (size_t i = 0; i < n; ++i) {
        // timing only (fmmfft_profile_end reads timestamps, no data): no
        // system-scope fence at each record, which would write back and
        // invalidate the caches between the stages it brackets
        hipError_t e = hipEventCreateWithFlags(&plan->prof_ev[i], hipEventDisableSystemFence);
        if (e != hipSuccess) {
            for (size_t j = 0; j < i; ++j) (void)hipEventDestroy(plan->prof_ev[j]);
            delete[] plan->prof_ev;
            plan->prof_ev = nullptr;
            g_last_hip_error.store((int)e);
            return FMMFFT_EHIP;
        }
    }
    plan->prof_cap = max_calls;
    plan->prof_calls = 0;
    return FMMFFT_OK;
}

int fmmfft_profile_end(fmmfft_plan plan, double *stage_ms, int32_t nstages, int64_t *calls) {
    if (!plan || !plan->prof_ev) return FMMFFT_EINVAL;
    const int ns = std::min<int32_t>(nstages, FMMFFT_NSTAGES);
    for (int i = 0; stage_ms && i < ns; ++i) stage_ms[i] = 0.0;
    int rc = FMMFFT_OK;
    for (int64_t c = 0; c < plan->prof_calls; ++c) {
        hipEvent_t *ev = plan->prof_ev + (size_t)c * (FMMFFT_NSTAGES + 1);
        hipError_t e = hipEventSynchronize(ev[FMMFFT_NSTAGES]);
        if (e != hipSuccess) {
            g_last_hip_error.store((int)e);
            rc = FMMFFT_EHIP;
            break;
        }
        for (int i = 0; stage_ms && i < ns; ++i) {
            float ms = 0.f;
            if (i == 6 && !plan->fft.two_pass) continue;  // single-pass FFT: stage 6 is empty
            if (hipEventElapsedTime(&ms, ev[i], ev[i + 1]) == hipSuccess) stage_ms[i] += ms;
        }
    }
    if (calls) *calls = plan->prof_calls;
    prof_free(plan);
    return rc;
}

int fmmfft_plan_destroy(fmmfft_plan plan) {
    if (!plan) return FMMFFT_EINVAL;
    prof_free(plan);
    if (plan->tab64) (void)hipFree(plan->tab64);
    if (plan->tab32) (void)hipFree(plan->tab32);
    if (plan->kfrag) (void)hipFree(plan->kfrag);
    if (plan->work) (void)hipFree(plan->work);
    delete plan;
    return FMMFFT_OK;
}

size_t fmmfft_workspace_bytes(fmmfft_plan plan, int64_t count) {
    if (!plan || count <= 0) return 0;
    return plan->ws.stride * (size_t)count;
}

int fmmfft_execute_ws(fmmfft_plan plan, int64_t count, const void *d_in, void *d_out, int32_t dir,
                      void *d_work, size_t work_bytes, void *stream) {
    if (!plan || !d_in || !d_out || !d_work || count <= 0 || count > 65535) return FMMFFT_EINVAL;
    if (dir != FMMFFT_FORWARD && dir != FMMFFT_INVERSE) return FMMFFT_EINVAL;
    if (work_bytes < plan->ws.stride * (size_t)count || ((uintptr_t)d_work % kAlign) != 0) return FMMFFT_EINVAL;
    // the S2M kernel streams the input with 16-byte loads
    if (((uintptr_t)d_in % 16) != 0 || ((uintptr_t)d_out % (2 * plan->prm.esize)) != 0) return FMMFFT_EINVAL;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t n = plan->prm.g.N;
    if (is_c64(plan->prm.dtype))
        return run<float, false>(plan, count, SigAddr<const c32>{static_cast<const c32 *>(d_in), n, nullptr},
                                 SigAddr<c32>{static_cast<c32 *>(d_out), n, nullptr}, dir, (char *)d_work, st);
    return run<double, false>(plan, count, SigAddr<const c64>{static_cast<const c64 *>(d_in), n, nullptr},
                              SigAddr<c64>{static_cast<c64 *>(d_out), n, nullptr}, dir, (char *)d_work, st);
}

}  // extern "C"
#endif  // !FMMFFT_TAB_UNIT

#ifdef FMMFFT_TAB_UNIT
int fmmfft::execute_table(fmmfft_plan plan, int64_t count, const void *const *d_in_tab, void *const *d_out_tab,
                          int32_t dir, void *d_work, size_t work_bytes, void *stream) {
    if (!plan || !d_in_tab || !d_out_tab || !d_work || count <= 0 || count > 65535) return FMMFFT_EINVAL;
    if (dir != FMMFFT_FORWARD && dir != FMMFFT_INVERSE) return FMMFFT_EINVAL;
    if (work_bytes < plan->ws.stride * (size_t)count || ((uintptr_t)d_work % kAlign) != 0) return FMMFFT_EINVAL;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (is_c64(plan->prm.dtype))
        return run<float, true>(plan, count, SigAddr<const c32>{nullptr, 0, reinterpret_cast<const c32 *const *>(d_in_tab)},
                                SigAddr<c32>{nullptr, 0, reinterpret_cast<c32 *const *>(d_out_tab)}, dir,
                                (char *)d_work, st);
    return run<double, true>(plan, count, SigAddr<const c64>{nullptr, 0, reinterpret_cast<const c64 *const *>(d_in_tab)},
                             SigAddr<c64>{nullptr, 0, reinterpret_cast<c64 *const *>(d_out_tab)}, dir,
                             (char *)d_work, st);
}
#else

extern "C" {

int fmmfft_execute_batched(fmmfft_plan plan, int64_t count, const void *d_in, void *d_out, int32_t dir,
                           void *stream) {
    if (!plan || !plan->work || count <= 0 || count > plan->max_batch) return FMMFFT_EINVAL;
    return fmmfft_execute_ws(plan, count, d_in, d_out, dir, plan->work, plan->ws.stride * (size_t)count,
                             stream);
}

int fmmfft_execute(fmmfft_plan plan, const void *d_in, void *d_out, int32_t dir, void *stream) {
    return fmmfft_execute_batched(plan, 1, d_in, d_out, dir, stream);
}

int fmmfft_execute_static(fmmfft_plan plan, const void *d_in, void *d_out, int32_t dir, void *stream) {
    if (!plan) return FMMFFT_EINVAL;
    if (plan->ws.stride > (size_t)FMMFFT_STATIC_SCRATCH_BYTES) return FMMFFT_ETOOBIG;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    void *ws = static_scratch(dev);
    if (!ws) return FMMFFT_EHIP;
    return fmmfft_execute_ws(plan, 1, d_in, d_out, dir, ws, plan->ws.stride, stream);
}

int fmmfft_plan_get(fmmfft_plan plan, int64_t *n, int32_t *p_blocks, int32_t *terms, int32_t *leaf,
                    int32_t *dtype, int32_t *depth) {
    if (!plan) return FMMFFT_EINVAL;
    const Geom &g = plan->prm.g;
    if (n) *n = g.N;
    if (p_blocks) *p_blocks = g.P;
    if (terms) *terms = g.terms;
    if (leaf) *leaf = g.b;
    if (dtype) *dtype = plan->prm.dtype;
    if (depth) *depth = g.n;
    return FMMFFT_OK;
}

int fmmfft_read_table(fmmfft_plan plan, int32_t which, void *host_dst, size_t bytes) {
    if (!plan || !host_dst) return FMMFFT_EINVAL;
    const Geom &g = plan->prm.g;
    const TableLayout &tl = plan->tl;
    size_t off, cnt;
    switch (which) {
        case 0: off = tl.cheb; cnt = g.terms; break;
        case 1: off = tl.initch; cnt = (size_t)g.b * g.terms; break;
        case 2: off = tl.ops; cnt = (size_t)(g.n - 2) * NOPS * g.terms * g.terms; break;
        case 3: off = tl.topflp; cnt = (size_t)g.terms * g.terms; break;
        case 4: off = tl.evalm; cnt = (size_t)g.S * g.b * g.terms; break;
        case 5: off = tl.evalmh; cnt = (size_t)(g.P / 2) * g.terms; break;
        case 6: off = tl.kv; cnt = (size_t)g.S * 4 * g.b; break;
        case 7: off = tl.gfac; cnt = 2 * (size_t)g.P; break;
        default: return FMMFFT_EINVAL;
    }
    if (bytes != cnt * sizeof(double)) return FMMFFT_EINVAL;
    HIP_TRY(hipMemcpy(host_dst, plan->tab64 + off, bytes, hipMemcpyDeviceToHost));
    return FMMFFT_OK;
}

int fmmfft_read_stage(fmmfft_plan plan, int32_t which, void *host_dst, size_t bytes) {
    if (!plan || !host_dst || !plan->work) return FMMFFT_EINVAL;
    const Geom &g = plan->prm.g;
    const WsLayout &wl = plan->ws;
    size_t off, cnt;
    switch (which) {
        case 0: off = wl.sumsec; cnt = (size_t)g.P * 16; break;
        case 1: off = wl.phi; cnt = (size_t)tree_len(g) * plan->prm.esize; break;
        case 2: off = wl.psi; cnt = (size_t)tree_len(g) * plan->prm.esize; break;
        case 3: off = wl.psec; cnt = (size_t)g.N * 2 * plan->prm.esize; break;
        default: return FMMFFT_EINVAL;
    }
    if (which == 1 || which == 2) {
        // the tree as [node][terms][S][2]: the rows of section 0 are dropped
        const size_t es = plan->prm.esize, nrows = (size_t)tree_len(g) / tree_rows(g);
        if (bytes != nrows * 2 * (size_t)g.S * es) return FMMFFT_EINVAL;
        HIP_TRY(hipDeviceSynchronize());
        HIP_TRY(hipMemcpy2D(host_dst, 2 * (size_t)g.S * es, (char *)plan->work + off + 2 * es,
                            (size_t)tree_rows(g) * es, 2 * (size_t)g.S * es, nrows, hipMemcpyDeviceToHost));
        return FMMFFT_OK;
    }
    if (bytes != cnt) return FMMFFT_EINVAL;
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(host_dst, (char *)plan->work + off, bytes, hipMemcpyDeviceToHost));
    return FMMFFT_OK;
}

const char *fmmfft_strerror(int status) {
    switch (status) {
        case FMMFFT_OK: return "ok";
        case FMMFFT_EINVAL: return "invalid argument";
        case FMMFFT_ENOMEM: return "out of device memory";
        case FMMFFT_EHIP: return "HIP runtime error";
        case FMMFFT_ENOTSUP: return "parameters not supported by the kernels";
        case FMMFFT_ETOOBIG: return "workspace larger than the static scratch (FMMFFT_STATIC_SCRATCH_BYTES)";
        default: return "unknown status";
    }
}

int fmmfft_last_hip_error(void) { return g_last_hip_error.load(); }

}  // extern "C"
#endif  // FMMFFT_TAB_UNIT
