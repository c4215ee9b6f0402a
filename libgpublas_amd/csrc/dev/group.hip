// group.hip -- grouped persistent launches for the serving loop: one batch
// slot holds up to 24 GEMMs per tile shape and precision, 32 DGEMVs and 48
// DTRSMs of arbitrary, uneven sizes and runs them all with one replay of a
// captured hipGraph (or the same kernels launched eagerly).
//
// No counterpart in the reference: blas2cuda issues one cuBLAS call per
// intercepted BLAS call (blas_level3/gemm.cc:62-82, trsm.cc:52-72) on one
// global handle.  Here the host engine collects calls from many threads and
// hands them over in batches.
//
// Problem descriptors travel per kind in a group struct (at most 4 KB):
// eager launches pass each group by value as the kernel argument; a graph
// replay reads the same structs from a device copy of the batch table,
// refreshed by the graph's first node.  Besides the operands themselves
// (device-resident tracked objects, addressed in place) a batch uses a
// GEMV partial-sum workspace in HBM.  Completion is an event recorded after
// the launch.
//
// Kernels (all persistent: a fixed grid loops over the batch's work units,
// so one captured graph serves every batch of the slot):
//  * group_gemm_arg / _tab<T, BM>: every tile of every GEMM problem that the
//    tile-shape table assigns to BM x BM, full tiles first, in one grid;
//    tile t -> problem by binary search over the prefix sums; the tile
//    itself is the per-call fast tile routine (gemm_tile.hpp; its guarded
//    variant for ragged tiles), so a product's bits do not depend on the
//    batch;
//  * group_gemv_part_kernel / group_gemv_fold_kernel: y = alpha*op(A)*x +
//    beta*y as column-block (N) or row-chunk (T) partial sums followed by a
//    fixed-order fold; the split depends only on (m, n);
//  * group_trsm_kernel: left-side triangular solves (m <= 1024), one
//    workgroup per 16-column panel of B held in LDS, right-looking over
//    16-row blocks (in-register substitution of the diagonal block, MFMA
//    updates of the blocks below);
// Roofline: GEMM tiles are MFMA-bound (2mnk flop), GEMV is HBM-bound
// (8(mn + m + n) bytes), TRSM panels are bound by their sequential
// substitution chain for small n.
#include "gemm_tile.hpp"
#include "gpublas_dev.h"

#include <mutex>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

namespace {

enum { CFG_32 = 0, CFG_64 = 1, CFG_128 = 2, kCfgs = 3 };
enum { TY_D = 0, TY_S = 1 };
// problems per kind per batch, sized so each group struct stays a legal
// (<= 4 KB) kernel argument
constexpr int kCapGemm = 22, kCapGemv = 32, kCapTrsm = 48;

// One GEMM of a group.  Its tiles split into interior tiles (g.full_m x
// g.full_n, only when both operands allow the vector loads), run by the
// fast kernel with the trans pattern `pat`, and edge tiles (the rest), run
// by the generic kernel.  A product with few tiles and a long k is also
// split along k into `ksplit` chunks of `kchunk` (a multiple of 16): chunk
// s computes its tiles into part + s*m*n (alpha 1, beta 0) and the fold
// kernel sums the chunks in order and applies alpha / beta.  The split
// depends only on (m, n, k).
template <typename T>
struct GGemm {
    GemmArgs<T> g;
    T *part;
    int fast_begin, edge_begin, fold_begin;
    int pat;  // 0 NN, 1 NT, 2 TN, 3 TT; + 4: an operand not aligned (element loads)
    int ksplit, kchunk;
};

struct GGemv {
    const double *A;
    const double *x;
    double *y;
    double *part;  // split x len partial sums
    long lda;
    int incx, incy;
    int m, n, trans;
    int split;   // N: column blocks; T: row chunks
    int span;    // N: columns per block; T: rows per chunk
    int vec_ok;  // A 16-B aligned, lda even
    int x_vec;   // x 16-B aligned, incx == 1
    int direct;  // split == 1: the partial kernel writes y itself
    long x0, y0; // offsets of logical element 0 (negative strides start at the far end)
    double alpha, beta;
    int unit_begin, fold_begin;
};

struct GTrsm {
    const double *A;
    double *B;
    long lda, ldb;
    int m, n;
    int fwd;    // op(A) lower: forward substitution
    int trans;  // op(A) = A^T
    int unit;
    double alpha;
    int unit_begin;  // panels
};

template <typename T>
struct GemmGroup {
    int np, fast_tiles, edge_tiles, fold_units;
    GGemm<T> p[kCapGemm];
};
struct GemvGroup {
    int np, units, fold_units;
    GGemv p[kCapGemv];
};
struct TrsmGroup {
    int np, units;
    GTrsm p[kCapTrsm];
};
static_assert(sizeof(GemmGroup<double>) <= 4000 && sizeof(GemvGroup) <= 4000 &&
                  sizeof(TrsmGroup) <= 4000,
              "group structs must fit a kernel argument");

struct BatchTable {
    GemmGroup<double> dg[kCfgs];
    GemmGroup<float> sg[kCfgs];
    GemvGroup gv;
    TrsmGroup tr;
};

__device__ __forceinline__ int find_unit(const int *begin, size_t stride_bytes, int n, int u)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int b = *(const int *)((const char *)begin + (size_t)mid * stride_bytes);
        if (b <= u) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ---- grouped GEMM -----------------------------------------------------------
// The product restricted to k-chunk s of a split GEMM (identity if unsplit).
template <typename T>
__device__ __forceinline__ GemmArgs<T> chunk_args(const GGemm<T> &e, int s)
{
    GemmArgs<T> g = e.g;
    if (e.ksplit == 1) return g;
    const long k0 = (long)s * e.kchunk;
    g.k = min(e.kchunk, e.g.k - (int)k0);
    g.a.p += g.a.form == PAT_OUTER ? k0 * g.a.ld : k0;
    g.b.p += g.b.form == PAT_OUTER ? k0 * g.b.ld : k0;
    g.C = e.part + (long)s * g.m * g.n;
    g.ldc = g.m;
    g.alpha = T(1);
    g.beta = T(0);
    return g;
}

// Full tile t of the group (numbered over its problems' fast tiles).
template <typename T, int BM>
__device__ __forceinline__ void fast_group_tile(const GemmGroup<T> &G, int np, int t,
                                                T (*As)[16][BM + 128 / sizeof(T)],
                                                T (*Bs)[16][BM + 128 / sizeof(T)])
{
    const int p = find_unit(&G.p[0].fast_begin, sizeof(GGemm<T>), np, t);
    const int per = G.p[p].g.full_m * G.p[p].g.full_n;
    const int tl0 = t - G.p[p].fast_begin, tl = tl0 % per;
    const GemmArgs<T> g = chunk_args(G.p[p], tl0 / per);
    const int tm = tl % g.full_m, tn = tl / g.full_m;
    switch (G.p[p].pat) {
    case 0: gemm_fast_tile<T, BM, BM, 16, PAT_OUTER, PAT_KCONT, SYM_NONE>(g, tm, tn, As, Bs); break;
    case 1: gemm_fast_tile<T, BM, BM, 16, PAT_OUTER, PAT_OUTER, SYM_NONE>(g, tm, tn, As, Bs); break;
    case 2: gemm_fast_tile<T, BM, BM, 16, PAT_KCONT, PAT_KCONT, SYM_NONE>(g, tm, tn, As, Bs); break;
    case 3: gemm_fast_tile<T, BM, BM, 16, PAT_KCONT, PAT_OUTER, SYM_NONE>(g, tm, tn, As, Bs); break;
    // operands not both aligned: element loads
    case 4: gemm_fast_tile<T, BM, BM, 16, PAT_OUTER_U, PAT_KCONT_U, SYM_NONE>(g, tm, tn, As, Bs); break;
    case 5: gemm_fast_tile<T, BM, BM, 16, PAT_OUTER_U, PAT_OUTER_U, SYM_NONE>(g, tm, tn, As, Bs); break;
    case 6: gemm_fast_tile<T, BM, BM, 16, PAT_KCONT_U, PAT_KCONT_U, SYM_NONE>(g, tm, tn, As, Bs); break;
    default: gemm_fast_tile<T, BM, BM, 16, PAT_KCONT_U, PAT_OUTER_U, SYM_NONE>(g, tm, tn, As, Bs); break;
    }
}

// Ragged tile t of the group (numbered over its problems' edge tiles; a
// problem's are the column strip past the full columns, all rows, followed
// by the row strip past the full rows), on the guarded fast tile with the
// problem's access patterns (element loads; same bits as the generic tile).
template <typename T, int BM>
__device__ __forceinline__ void edge_group_tile(const GemmGroup<T> &G, int np, int t,
                                                T (*As)[16][BM + 128 / sizeof(T)],
                                                T (*Bs)[16][BM + 128 / sizeof(T)])
{
    const int p = find_unit(&G.p[0].edge_begin, sizeof(GGemm<T>), np, t);
    const int per = G.p[p].g.tiles_m * G.p[p].g.tiles_n - G.p[p].g.full_m * G.p[p].g.full_n;
    const int e0 = t - G.p[p].edge_begin, e = e0 % per;
    const GemmArgs<T> g = chunk_args(G.p[p], e0 / per);
    const int strip = g.tiles_m * (g.tiles_n - g.full_n);
    int tm, tn;
    if (e < strip) {
        tm = e % g.tiles_m;
        tn = g.full_n + e / g.tiles_m;
    } else {
        const int e2 = e - strip, rows = g.tiles_m - g.full_m;
        tm = g.full_m + e2 % rows;
        tn = e2 / rows;
    }
    switch (G.p[p].pat & 3) {
    case 0: gemm_fast_tile<T, BM, BM, 16, PAT_OUTER, PAT_KCONT, SYM_NONE, true>(g, tm, tn, As, Bs); break;
    case 1: gemm_fast_tile<T, BM, BM, 16, PAT_OUTER, PAT_OUTER, SYM_NONE, true>(g, tm, tn, As, Bs); break;
    case 2: gemm_fast_tile<T, BM, BM, 16, PAT_KCONT, PAT_KCONT, SYM_NONE, true>(g, tm, tn, As, Bs); break;
    default: gemm_fast_tile<T, BM, BM, 16, PAT_KCONT, PAT_OUTER, SYM_NONE, true>(g, tm, tn, As, Bs); break;
    }
}

// Every tile of the group, full ones first, in one grid: the ragged tiles
// of a batch do not make a round of their own after the full ones.
template <typename T, int BM>
__device__ __forceinline__ void gemm_group_body(const GemmGroup<T> &G)
{
    constexpr int PAD = 128 / (int)sizeof(T);
    __shared__ __attribute__((aligned(16))) T As[2][16][BM + PAD];
    __shared__ __attribute__((aligned(16))) T Bs[2][16][BM + PAD];
    const int np = min(G.np, kCapGemm);
    const int nf = BM == 32 ? 0 : G.fast_tiles, total = nf + G.edge_tiles;
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        if (t < nf) {
            if constexpr (BM != 32) fast_group_tile<T, BM>(G, np, t, As, Bs);
        } else {
            edge_group_tile<T, BM>(G, np, t - nf, As, Bs);
        }
    }
}

// C = alpha * sum_s part_s + beta * C for the split products, 256 elements
// per unit, chunks summed in order.
template <typename T>
__device__ __forceinline__ void gemm_fold_body(const GemmGroup<T> &G)
{
    const int np = min(G.np, kCapGemm);
    const int total = G.fold_units;
    for (int u = blockIdx.x; u < total; u += gridDim.x) {
        const int p = find_unit(&G.p[0].fold_begin, sizeof(GGemm<T>), np, u);
        const GGemm<T> &e = G.p[p];
        const long mn = (long)e.g.m * e.g.n;
        const long idx = (long)(u - e.fold_begin) * 256 + threadIdx.x;
        if (idx >= mn) continue;
        T s = T(0);
        for (int q = 0; q < e.ksplit; ++q) s += e.part[(long)q * mn + idx];
        T *c = e.g.C + idx % e.g.m + (idx / e.g.m) * e.g.ldc;
        const T v = e.g.alpha * s;
        *c = e.g.beta == T(0) ? v : v + e.g.beta * *c;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void group_gemm_fold_arg(const GemmGroup<T> G)
{
    gemm_fold_body<T>(G);
}
template <typename T>
__global__ __launch_bounds__(256) void group_gemm_fold_tab(const GemmGroup<T> *G)
{
    gemm_fold_body<T>(*G);
}

template <typename T, int BM>
__global__ __launch_bounds__(256, 2) void group_gemm_arg(const GemmGroup<T> G)
{
    gemm_group_body<T, BM>(G);
}
template <typename T, int BM>
__global__ __launch_bounds__(256, 2) void group_gemm_tab(const GemmGroup<T> *G)
{
    gemm_group_body<T, BM>(*G);
}

// ---- grouped GEMV -----------------------------------------------------------
constexpr int GV_ROWS = 512;    // N: rows per unit (2 per thread)
constexpr int GV_TCOLS = 16;    // T: columns per unit (4 per wave)
constexpr int GV_FOLD = 256;    // fold: outputs per unit

__device__ __forceinline__ void gemv_out(const GGemv &d, long i, double acc)
{
    double *y = d.y + d.y0 + i * d.incy;
    const double v = d.alpha * acc;
    *y = d.beta == 0.0 ? v : v + d.beta * *y;
}

// N: rows [rb*512, +512) (2 per thread, wave w at rows 128w..) over the
// columns [cb*span, +span), eight 16-B column loads in flight per lane
__device__ void gemv_n_unit(const GGemv &d, int local)
{
    typedef double d2 __attribute__((ext_vector_type(2)));
    const int nrb = (d.m + GV_ROWS - 1) / GV_ROWS;
    const int rb = local % nrb, cb = local / nrb;
    const long r0 = (long)rb * GV_ROWS + 2 * (int)threadIdx.x;
    const int c0 = cb * d.span, c1 = min(d.n, c0 + d.span);
    const double *x = d.x + d.x0;
    const long incx = d.incx;
    double acc0 = 0.0, acc1 = 0.0;
    if (r0 >= d.m) return;
    const bool pair = r0 + 1 < d.m;
    if (pair && d.vec_ok) {
        const double *a = d.A + r0;
        int j = c0;
        for (; j + 8 <= c1; j += 8) {
            d2 v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                v[u] = __builtin_nontemporal_load((const d2 *)(a + (long)(j + u) * d.lda));
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const double xv = x[(long)(j + u) * incx];
                acc0 = fma(v[u].x, xv, acc0);
                acc1 = fma(v[u].y, xv, acc1);
            }
        }
        for (; j < c1; ++j) {
            const d2 v = *(const d2 *)(a + (long)j * d.lda);
            const double xv = x[(long)j * incx];
            acc0 = fma(v.x, xv, acc0);
            acc1 = fma(v.y, xv, acc1);
        }
    } else {
        for (int j = c0; j < c1; ++j) {
            const double xv = x[(long)j * incx];
            const double *a = d.A + r0 + (long)j * d.lda;
            acc0 = fma(a[0], xv, acc0);
            if (pair) acc1 = fma(a[1], xv, acc1);
        }
    }
    if (d.direct) {
        gemv_out(d, r0, acc0);
        if (pair) gemv_out(d, r0 + 1, acc1);
        return;
    }
    double *p = d.part + (long)cb * d.m + r0;
    p[0] = acc0;
    if (pair) p[1] = acc1;
}

// T: partial dot products of 16 columns over rows [rc*span, +span); wave w
// owns columns 4w..4w+3, each lane takes row pairs 128 apart and keeps four
// row steps (16 loads) in flight
__device__ void gemv_t_unit(const GGemv &d, int local)
{
    typedef double d2 __attribute__((ext_vector_type(2)));
    const int ncb = (d.n + GV_TCOLS - 1) / GV_TCOLS;
    const int cb = local % ncb, rc = local / ncb;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j0 = cb * GV_TCOLS + wave * 4;
    const int r0 = rc * d.span, r1 = min(d.m, r0 + d.span);
    const double *x = d.x + d.x0;
    const long incx = d.incx;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const double *a[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) a[c] = d.A + (long)min(j0 + c, d.n - 1) * d.lda;
    const bool cok[4] = {j0 < d.n, j0 + 1 < d.n, j0 + 2 < d.n, j0 + 3 < d.n};
    int i = r0 + 2 * lane;
    if (d.vec_ok && d.x_vec) {
        // full steps: 4 x 128 rows, every row pair inside [r0, r1)
        for (; i + 3 * 128 + 1 < r1; i += 4 * 128) {
            d2 av[4][4], xv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                xv[q] = *(const d2 *)(x + i + 128 * q);
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    av[q][c] = cok[c] ? __builtin_nontemporal_load((const d2 *)(a[c] + i + 128 * q))
                                      : d2{0.0, 0.0};
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    acc[c] = fma(av[q][c].x, xv[q].x, acc[c]);
                    acc[c] = fma(av[q][c].y, xv[q].y, acc[c]);
                }
        }
    }
    // remaining rows (and the general-stride / unaligned case)
    for (; i < r1; i += 128) {
        const bool pair = i + 1 < r1;
        const double x0 = x[(long)i * incx];
        const double x1 = pair ? x[(long)(i + 1) * incx] : 0.0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if (!cok[c]) continue;
            acc[c] = fma(a[c][i], x0, acc[c]);
            if (pair) acc[c] = fma(a[c][i + 1], x1, acc[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        double v = acc[c];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0 && cok[c]) {
            if (d.direct) gemv_out(d, j0 + c, v);
            else d.part[(long)rc * d.n + j0 + c] = v;
        }
    }
}

__device__ __forceinline__ void gemv_part_body(const GemvGroup &G)
{
    const int np = min(G.np, kCapGemv);
    const int total = G.units;
    for (int u = blockIdx.x; u < total; u += gridDim.x) {
        const int p = find_unit(&G.p[0].unit_begin, sizeof(GGemv), np, u);
        const GGemv d = G.p[p];
        if (d.trans) gemv_t_unit(d, u - d.unit_begin);
        else gemv_n_unit(d, u - d.unit_begin);
    }
}

__device__ __forceinline__ void gemv_fold_body(const GemvGroup &G)
{
    const int np = min(G.np, kCapGemv);
    const int total = G.fold_units;
    for (int u = blockIdx.x; u < total; u += gridDim.x) {
        const int p = find_unit(&G.p[0].fold_begin, sizeof(GGemv), np, u);
        const GGemv d = G.p[p];
        const int len = d.trans ? d.n : d.m;
        const int i = (u - d.fold_begin) * GV_FOLD + (int)threadIdx.x;
        if (i >= len) continue;
        const double *pp = d.part + i;
        double s = 0.0;
        int q = 0;
        for (; q + 4 <= d.split; q += 4) {
            const double p0 = pp[(long)q * len], p1 = pp[(long)(q + 1) * len];
            const double p2 = pp[(long)(q + 2) * len], p3 = pp[(long)(q + 3) * len];
            s += p0;
            s += p1;
            s += p2;
            s += p3;
        }
        for (; q < d.split; ++q) s += pp[(long)q * len];
        gemv_out(d, i, s);
    }
}

__global__ __launch_bounds__(256) void group_gemv_part_arg(const GemvGroup G) { gemv_part_body(G); }
__global__ __launch_bounds__(256) void group_gemv_part_tab(const GemvGroup *G) { gemv_part_body(*G); }
__global__ __launch_bounds__(256) void group_gemv_fold_arg(const GemvGroup G) { gemv_fold_body(G); }
__global__ __launch_bounds__(256) void group_gemv_fold_tab(const GemvGroup *G) { gemv_fold_body(*G); }

// ---- grouped TRSM (left side) ----------------------------------------------
// One 1024-thread workgroup solves op(A) X = alpha B for a 16-column panel
// of B with the panel resident in LDS (m <= 1024), right-looking over
// 16-row blocks: the owner wave of block I substitutes X_I in registers
// from the diagonal block, then every wave subtracts op(A)[J, I] X_I from
// the blocks J it owns (J = wave + 16 s) with one MFMA chain per block.
// The op(A) column blocks of the next step are loaded into registers while
// the current step runs.  Subtractions reach each row of B in increasing
// (forward) or decreasing (backward) block order, as in the netlib loop.
constexpr int TR_MAXM = 1024;
constexpr int TR_WAVES = 16;

__device__ __forceinline__ double opa(const GTrsm &d, int i, int k)
{
    if (i >= d.m || k >= d.m) return 0.0;
    return d.trans ? d.A[k + (long)i * d.lda] : d.A[i + (long)k * d.lda];
}

// op(A)[J rows, I cols] for the blocks J = wave + 16 s this wave will touch
// at step I (J >= I forward, J <= I backward), in MFMA A-operand layout:
// buf[s][u] = op(A)(16 J + l%16, 16 I + 4 u + l/16).
__device__ __forceinline__ void trsm_load_col(const GTrsm &d, int I, int nb, int wave, int q0,
                                              int cl, double (&buf)[4][4])
{
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int J = wave + TR_WAVES * s;
        const bool need = J < nb && (d.fwd ? J >= I : J <= I);
#pragma unroll
        for (int u = 0; u < 4; ++u)
            buf[s][u] = need ? opa(d, 16 * J + cl, 16 * I + 4 * u + q0) : 0.0;
    }
}

__device__ __forceinline__ void trsm_group_body(const TrsmGroup &G)
{
    __shared__ double Bs[TR_MAXM][17];
    __shared__ double Ds[16][17];
    const int np = min(G.np, kCapTrsm);
    const int total = G.units;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int q0 = lane >> 4, cl = lane & 15;
    for (int u = blockIdx.x; u < total; u += gridDim.x) {
        const int p = find_unit(&G.p[0].unit_begin, sizeof(GTrsm), np, u);
        const GTrsm d = G.p[p];
        const int c0 = (u - d.unit_begin) * 16;
        const int nb = (d.m + 15) / 16;
        const int rows = nb * 16;
        // alpha * B panel into LDS, zero past m and past n
        for (int idx = threadIdx.x; idx < rows * 16; idx += 1024) {
            const int r = idx % rows, c = idx / rows;
            double v = 0.0;
            if (r < d.m && c0 + c < d.n) v = d.alpha * d.B[r + (long)(c0 + c) * d.ldb];
            Bs[r][c] = v;
        }
        double cur[4][4], nxt[4][4];
        trsm_load_col(d, d.fwd ? 0 : nb - 1, nb, wave, q0, cl, cur);
        __syncthreads();
        for (int st = 0; st < nb; ++st) {
            const int I = d.fwd ? st : nb - 1 - st;
            const int I0 = 16 * I;
            if (st + 1 < nb) trsm_load_col(d, d.fwd ? I + 1 : I - 1, nb, wave, q0, cl, nxt);
            if (wave == (I & (TR_WAVES - 1))) {
                // diagonal block (identity past m; unit diagonal if asked)
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    if (s != I / TR_WAVES) continue;
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int k = 4 * v + q0;
                        double a = cur[s][v];
                        if (k == cl && (d.unit || I0 + cl >= d.m)) a = 1.0;
                        Ds[cl][k] = a;
                    }
                }
                double t[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) t[e] = Bs[I0 + q0 + 4 * e][cl];
                asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
                __builtin_amdgcn_wave_barrier();
                if (d.fwd) {
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int q = r & 3, e = r >> 2;
                        const double v = t[e] / Ds[r][r];
                        const double xr = __shfl(v, cl + 16 * q);
                        if (q0 == q) t[e] = xr;
#pragma unroll
                        for (int e2 = 0; e2 < 4; ++e2)
                            if (q0 + 4 * e2 > r) t[e2] = fma(-Ds[q0 + 4 * e2][r], xr, t[e2]);
                    }
                } else {
#pragma unroll
                    for (int r = 15; r >= 0; --r) {
                        const int q = r & 3, e = r >> 2;
                        const double v = t[e] / Ds[r][r];
                        const double xr = __shfl(v, cl + 16 * q);
                        if (q0 == q) t[e] = xr;
#pragma unroll
                        for (int e2 = 0; e2 < 4; ++e2)
                            if (q0 + 4 * e2 < r) t[e2] = fma(-Ds[q0 + 4 * e2][r], xr, t[e2]);
                    }
                }
                const bool cok = c0 + cl < d.n;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int row = I0 + q0 + 4 * e;
                    Bs[row][cl] = t[e];
                    if (cok && row < d.m) d.B[row + (long)(c0 + cl) * d.ldb] = t[e];
                }
            }
            __syncthreads();
            // B_J -= op(A)[J, I] X_I for this wave's remaining blocks
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int J = wave + TR_WAVES * s;
                if (J >= nb || (d.fwd ? J <= I : J >= I)) continue;
                const int J0 = 16 * J;
                v4d acc;
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[r] = Bs[J0 + q0 + 4 * r][cl];
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(-cur[s][v], Bs[I0 + 4 * v + q0][cl],
                                                              acc, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; ++r) Bs[J0 + q0 + 4 * r][cl] = acc[r];
            }
            __syncthreads();
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int v = 0; v < 4; ++v) cur[s][v] = nxt[s][v];
        }
    }
}

__global__ __launch_bounds__(1024) void group_trsm_arg(const TrsmGroup G) { trsm_group_body(G); }
__global__ __launch_bounds__(1024) void group_trsm_tab(const TrsmGroup *G) { trsm_group_body(*G); }

// ---- tile-shape table (autotuned) ------------------------------------------
// Bucket = (round(log2 sqrt(m n)), round(log2 k)), each clamped to 5..11.
constexpr int kB = 7;
std::mutex g_tune_mu;
signed char g_tune[2][kB][kB];  // -1 = heuristic
bool g_tune_init = false;

int ilog2_round(double v)
{
    int e = 0;
    while (v >= 1.4142135623730951 && e < 40) {
        v *= 0.5;
        ++e;
    }
    return e;
}
int bucket(double v)
{
    int e = ilog2_round(v);
    if (e < 5) e = 5;
    if (e > 11) e = 11;
    return e - 5;
}

bool cfg_valid(int ty, int cfg) { return !(ty == TY_S && cfg == CFG_32); }

int heuristic_cfg(int ty, int m, int n)
{
    if (m >= 384 && n >= 384) return CFG_128;
    if (ty == TY_D && m <= 48 && n <= 48) return CFG_32;
    return CFG_64;
}

void tune_init_locked()
{
    if (g_tune_init) return;
    memset(g_tune, -1, sizeof g_tune);
    g_tune_init = true;
}

int choose_cfg(int ty, int m, int n, int k)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_init_locked();
    const int c = g_tune[ty][bucket(sqrt((double)m * (double)n))][bucket((double)k)];
    return c >= 0 && cfg_valid(ty, c) ? c : heuristic_cfg(ty, m, n);
}

int cfg_bm(int cfg) { return cfg == CFG_32 ? 32 : cfg == CFG_64 ? 64 : 128; }

// grid of each persistent kernel: enough workgroups to fill 256 CUs at the
// kernel's occupancy (double / single), capped by the work when eager
constexpr int kGridGemm[2][kCfgs] = {{1024, 512, 512}, {1024, 512, 512}};  // 2 per CU from 64
constexpr int kGridGemv = 1024, kGridFold = 512, kGridTrsm = 256;  // TRSM: 1024-thread groups
constexpr size_t kWsDoubles = (size_t)8 << 20;  // 64 MiB of GEMV partial sums

}  // namespace

// ---- batch slot -------------------------------------------------------------
struct gpublas_batch {
    int use_graph;
    int force_cfg;
    hipStream_t s;
    hipStream_t br[3];
    hipEvent_t ev_fork, ev_join[3], ev_done;
    hipEvent_t ev_kind[4];  // eager: completion of each kind (DGEMM, SGEMM, DGEMV, DTRSM)
    int kind_used[4];
    BatchTable *h_tab;  // pinned; eager launches pass its groups by value
    BatchTable *d_tab;  // graph replays read this copy
    double *d_ws;
    size_t ws_cap, ws_used;  // doubles
    int launched;
    hipGraphExec_t exec[256];
    long graph_launches, eager_launches;
};

namespace {

void batch_clear(gpublas_batch *b)
{
    BatchTable *t = b->h_tab;
    for (int k = 0; k < 4; ++k) b->kind_used[k] = 0;
    for (int c = 0; c < kCfgs; ++c) {
        t->dg[c].np = t->dg[c].fast_tiles = t->dg[c].edge_tiles = t->dg[c].fold_units = 0;
        t->sg[c].np = t->sg[c].fast_tiles = t->sg[c].edge_tiles = t->sg[c].fold_units = 0;
    }
    t->gv.np = t->gv.units = t->gv.fold_units = 0;
    t->tr.np = t->tr.units = 0;
    b->ws_used = 0;
}

template <typename T> GemmGroup<T> &group_of(BatchTable *t, int cfg);
template <> GemmGroup<double> &group_of<double>(BatchTable *t, int cfg) { return t->dg[cfg]; }
template <> GemmGroup<float> &group_of<float>(BatchTable *t, int cfg) { return t->sg[cfg]; }

template <typename T>
int add_gemm(gpublas_batch *b, char ta, char tb, int m, int n, int k, T alpha, const T *A, int lda,
             const T *B, int ldb, T beta, T *C, int ldc)
{
    constexpr int ty = sizeof(T) == 8 ? TY_D : TY_S;
    if (m <= 0 || n <= 0 || k <= 0 || alpha == T(0)) return -1;
    const int cfg = b->force_cfg >= 0 && cfg_valid(ty, b->force_cfg) ? b->force_cfg
                                                                     : choose_cfg(ty, m, n, k);
    GemmGroup<T> &G = group_of<T>(b->h_tab, cfg);
    if (G.np >= kCapGemm) return 1;
    GGemm<T> &e = G.p[G.np];
    memset(&e, 0, sizeof e);
    GemmArgs<T> &g = e.g;
    g.m = m;
    g.n = n;
    g.k = k;
    g.alpha = alpha;
    g.beta = beta;
    g.a = make_op<T>(A, lda, m, is_n(ta) ? PAT_OUTER : PAT_KCONT, SYM_NONE);
    g.b = make_op<T>(B, ldb, n, is_n(tb) ? PAT_KCONT : PAT_OUTER, SYM_NONE);
    g.C = C;
    g.ldc = ldc;
    const int bm = cfg_bm(cfg);
    g.tiles_m = ceil_div(m, bm);
    g.tiles_n = ceil_div(n, bm);
    g.tri = SYM_NONE;
    // interior tiles take the fast tile: vector loads when both operands
    // allow them, element loads otherwise (the 32-wide shape has no fast
    // variant)
    const bool fast = cfg != CFG_32;
    g.full_m = fast ? m / bm : 0;
    g.full_n = fast ? n / bm : 0;
    if (g.full_m == 0 || g.full_n == 0) g.full_m = g.full_n = 0;
    e.pat = (g.a.form == PAT_OUTER ? 0 : 2) + (g.b.form == PAT_KCONT ? 0 : 1) +
            (g.a.vec_ok && g.b.vec_ok ? 0 : 4);
    // split along k when the tiles alone cannot spread over the CUs: about
    // 256 units, chunks of at least 256, partials of at most 16 MiB.  Counted
    // in 64 x 64 tiles whatever the tile shape, so every shape splits alike
    // (and gives the same bits).
    const long tiles = (long)ceil_div(m, 64) * ceil_div(n, 64);
    long ks = (256 + tiles - 1) / tiles;
    const long kmax = k / 256;
    const long pmax = (16l << 20) / ((long)m * n * (long)sizeof(T));
    if (ks > kmax) ks = kmax;
    if (ks > pmax) ks = pmax;
    if (ks < 1) ks = 1;
    e.kchunk = (int)(((k + ks - 1) / ks + 15) / 16 * 16);
    e.ksplit = (k + e.kchunk - 1) / e.kchunk;
    e.part = nullptr;
    e.fold_begin = G.fold_units;
    if (e.ksplit > 1) {
        const size_t bytes = (size_t)e.ksplit * m * n * sizeof(T);
        const size_t words = (bytes + 255) / 256 * 32;  // doubles, 256-B aligned
        if (b->ws_used + words > b->ws_cap) return 1;
        e.part = (T *)(b->d_ws + b->ws_used);
        b->ws_used += words;
        G.fold_units += (int)(((long)m * n + 255) / 256);
    }
    e.fast_begin = G.fast_tiles;
    e.edge_begin = G.edge_tiles;
    G.fast_tiles += g.full_m * g.full_n * e.ksplit;
    G.edge_tiles += (g.tiles_m * g.tiles_n - g.full_m * g.full_n) * e.ksplit;
    G.np++;
    return 0;
}

unsigned batch_mask(const BatchTable *t)
{
    unsigned mask = 0;
    for (int c = 0; c < kCfgs; ++c) {
        if (t->dg[c].np) mask |= 1u << c;
        if (t->sg[c].np) mask |= 1u << (3 + c);
    }
    if (t->gv.units) mask |= 1u << 6;
    if (t->tr.units) mask |= 1u << 7;
    return mask;
}

template <typename T, int BM>
void launch_gemm_group(gpublas_batch *b, hipStream_t s, int cfg, bool capture)
{
    constexpr int ty = sizeof(T) == 8 ? TY_D : TY_S;
    const GemmGroup<T> &G = group_of<T>(b->h_tab, cfg);
    const int fix = kGridGemm[ty][cfg];
    if (capture) {
        const GemmGroup<T> *dG = &group_of<T>(b->d_tab, cfg);
        hipLaunchKernelGGL((group_gemm_tab<T, BM>), dim3(fix), dim3(256), 0, s, dG);
        hipLaunchKernelGGL((group_gemm_fold_tab<T>), dim3(512), dim3(256), 0, s, dG);
        return;
    }
    const int tiles = (BM == 32 ? 0 : G.fast_tiles) + G.edge_tiles;
    if (tiles)
        hipLaunchKernelGGL((group_gemm_arg<T, BM>), dim3(tiles < fix ? tiles : fix), dim3(256), 0,
                           s, G);
    if (G.fold_units)
        hipLaunchKernelGGL((group_gemm_fold_arg<T>), dim3(G.fold_units < 512 ? G.fold_units : 512),
                           dim3(256), 0, s, G);
}

// Estimated device time of each kind's share of the batch (microseconds,
// sustained rates of the batch kernels): eager launches run the cheaper
// kinds first so their callers can leave before the expensive ones end.
void kind_costs(const BatchTable *t, double (&cost)[4])
{
    cost[0] = cost[1] = cost[2] = cost[3] = 0.0;
    for (int c = 0; c < kCfgs; ++c) {
        for (int i = 0; i < t->dg[c].np; ++i) {
            const GemmArgs<double> &g = t->dg[c].p[i].g;
            cost[0] += 2.0 * g.m * (double)g.n * g.k / 30e6;
        }
        for (int i = 0; i < t->sg[c].np; ++i) {
            const GemmArgs<float> &g = t->sg[c].p[i].g;
            cost[1] += 2.0 * g.m * (double)g.n * g.k / 60e6;
        }
    }
    for (int i = 0; i < t->gv.np; ++i) cost[2] += 8.0 * t->gv.p[i].m * (double)t->gv.p[i].n / 3e6;
    for (int i = 0; i < t->tr.np; ++i) cost[3] += 0.04 * t->tr.p[i].m + 1.0;
}

// Enqueue the batch's kernels.  Captured: fixed persistent grids reading
// the device table (copied by the first node), the three kinds on parallel
// branches.  Eager: grids capped by the work, groups passed by value, kinds
// in increasing estimated cost, each followed by its completion event.
void enqueue(gpublas_batch *b, unsigned mask, bool capture)
{
    const BatchTable *t = b->h_tab;
    hipStream_t s0 = b->s, sg = b->s, sv = b->s, st = b->s;
    if (!capture) {
        double cost[4];
        kind_costs(t, cost);
        const unsigned kmask[4] = {1u | 2u | 4u, 16u | 32u, 64u, 128u};
        int order[4] = {0, 1, 2, 3};
        for (int i = 1; i < 4; ++i)
            for (int j = i; j > 0 && cost[order[j]] < cost[order[j - 1]]; --j) {
                const int tmp = order[j];
                order[j] = order[j - 1];
                order[j - 1] = tmp;
            }
        for (int q = 0; q < 4; ++q) {
            const int k = order[q];
            b->kind_used[k] = (mask & kmask[k]) != 0;
            if (!b->kind_used[k]) continue;
            if (k == 0) {
                if (mask & 1u) launch_gemm_group<double, 32>(b, s0, CFG_32, false);
                if (mask & 2u) launch_gemm_group<double, 64>(b, s0, CFG_64, false);
                if (mask & 4u) launch_gemm_group<double, 128>(b, s0, CFG_128, false);
            } else if (k == 1) {
                if (mask & 16u) launch_gemm_group<float, 64>(b, s0, CFG_64, false);
                if (mask & 32u) launch_gemm_group<float, 128>(b, s0, CFG_128, false);
            } else if (k == 2) {
                const int gp = t->gv.units < kGridGemv ? t->gv.units : kGridGemv;
                const int gf = t->gv.fold_units < kGridFold ? t->gv.fold_units : kGridFold;
                hipLaunchKernelGGL(group_gemv_part_arg, dim3(gp), dim3(256), 0, s0, t->gv);
                if (gf) hipLaunchKernelGGL(group_gemv_fold_arg, dim3(gf), dim3(256), 0, s0, t->gv);
            } else {
                const int g = t->tr.units < kGridTrsm ? t->tr.units : kGridTrsm;
                hipLaunchKernelGGL(group_trsm_arg, dim3(g), dim3(1024), 0, s0, t->tr);
            }
            (void)hipEventRecord(b->ev_kind[k], s0);
        }
        return;
    }
    (void)hipMemcpyAsync(b->d_tab, b->h_tab, sizeof(BatchTable), hipMemcpyHostToDevice, s0);
    (void)hipEventRecord(b->ev_fork, s0);
    sg = b->br[0];
    sv = b->br[1];
    st = b->br[2];
    (void)hipStreamWaitEvent(sg, b->ev_fork, 0);
    (void)hipStreamWaitEvent(sv, b->ev_fork, 0);
    (void)hipStreamWaitEvent(st, b->ev_fork, 0);
    if (mask & 1u) launch_gemm_group<double, 32>(b, sg, CFG_32, true);
    if (mask & 2u) launch_gemm_group<double, 64>(b, sg, CFG_64, true);
    if (mask & 4u) launch_gemm_group<double, 128>(b, sg, CFG_128, true);
    if (mask & 16u) launch_gemm_group<float, 64>(b, sg, CFG_64, true);
    if (mask & 32u) launch_gemm_group<float, 128>(b, sg, CFG_128, true);
    if (mask & 64u) {
        hipLaunchKernelGGL(group_gemv_part_tab, dim3(kGridGemv), dim3(256), 0, sv, &b->d_tab->gv);
        hipLaunchKernelGGL(group_gemv_fold_tab, dim3(kGridFold), dim3(256), 0, sv, &b->d_tab->gv);
    }
    if (mask & 128u)
        hipLaunchKernelGGL(group_trsm_tab, dim3(kGridTrsm), dim3(1024), 0, st, &b->d_tab->tr);
    (void)hipEventRecord(b->ev_join[0], sg);
    (void)hipEventRecord(b->ev_join[1], sv);
    (void)hipEventRecord(b->ev_join[2], st);
    for (int i = 0; i < 3; ++i) (void)hipStreamWaitEvent(s0, b->ev_join[i], 0);
}

// Branch streams exist only for capture: created on the first capture, so
// the slots' launch streams, created together, spread over the hardware
// queues (streams take queues round-robin in creation order).
bool ensure_branches(gpublas_batch *b)
{
    if (b->br[0]) return true;
    for (int i = 0; i < 3; ++i)
        if (hipStreamCreateWithFlags(&b->br[i], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&b->ev_join[i], hipEventDisableTiming) != hipSuccess)
            return false;
    return hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) == hipSuccess;
}

hipGraphExec_t graph_for(gpublas_batch *b, unsigned mask)
{
    if (b->exec[mask]) return b->exec[mask];
    if (!ensure_branches(b)) return nullptr;
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
    enqueue(b, mask, true);
    if (hipStreamEndCapture(b->s, &g) != hipSuccess || !g) return nullptr;
    hipGraphExec_t e = nullptr;
    if (hipGraphInstantiate(&e, g, nullptr, nullptr, 0) != hipSuccess) e = nullptr;
    (void)hipGraphDestroy(g);
    b->exec[mask] = e;
    return e;
}

}  // namespace

extern "C" {

gpublas_batch *gpublas_dev_batch_create(int use_graph)
{
    if (gpublas_dev_init() != 0) return nullptr;
    gpublas_batch *b = (gpublas_batch *)calloc(1, sizeof(gpublas_batch));
    if (!b) return nullptr;
    b->use_graph = use_graph;
    b->force_cfg = -1;
    bool ok = hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&b->ev_done, hipEventDisableTiming) == hipSuccess;
    for (int k = 0; k < 4 && ok; ++k)
        ok = hipEventCreateWithFlags(&b->ev_kind[k], hipEventDisableTiming) == hipSuccess;
    ok = ok && hipHostMalloc((void **)&b->h_tab, sizeof(BatchTable), hipHostMallocDefault) == hipSuccess;
    ok = ok && (!use_graph || hipMalloc((void **)&b->d_tab, sizeof(BatchTable)) == hipSuccess);
    ok = ok && hipMalloc((void **)&b->d_ws, kWsDoubles * sizeof(double)) == hipSuccess;
    if (!ok) {
        gpublas_dev_batch_destroy(b);
        return nullptr;
    }
    memset(b->h_tab, 0, sizeof(BatchTable));
    b->ws_cap = kWsDoubles;
    batch_clear(b);
    return b;
}

void gpublas_dev_batch_destroy(gpublas_batch *b)
{
    if (!b) return;
    if (b->s) (void)hipStreamSynchronize(b->s);
    for (int i = 0; i < 256; ++i)
        if (b->exec[i]) (void)hipGraphExecDestroy(b->exec[i]);
    for (int i = 0; i < 3; ++i) {
        if (b->br[i]) (void)hipStreamDestroy(b->br[i]);
        if (b->ev_join[i]) (void)hipEventDestroy(b->ev_join[i]);
    }
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->ev_done) (void)hipEventDestroy(b->ev_done);
    for (int k = 0; k < 4; ++k)
        if (b->ev_kind[k]) (void)hipEventDestroy(b->ev_kind[k]);
    if (b->s) (void)hipStreamDestroy(b->s);
    if (b->h_tab) (void)hipHostFree(b->h_tab);
    if (b->d_tab) (void)hipFree(b->d_tab);
    if (b->d_ws) (void)hipFree(b->d_ws);
    free(b);
}

void gpublas_dev_batch_clear(gpublas_batch *b) { batch_clear(b); }

int gpublas_dev_batch_count(const gpublas_batch *b)
{
    const BatchTable *t = b->h_tab;
    int n = t->gv.np + t->tr.np;
    for (int c = 0; c < kCfgs; ++c) n += t->dg[c].np + t->sg[c].np;
    return n;
}

void gpublas_dev_batch_force_cfg(gpublas_batch *b, int cfg) { b->force_cfg = cfg; }

void *gpublas_dev_batch_stream(gpublas_batch *b) { return b->s; }

int gpublas_dev_batch_dgemm(gpublas_batch *b, char transa, char transb, int m, int n, int k,
                            double alpha, const double *d_A, int lda, const double *d_B, int ldb,
                            double beta, double *d_C, int ldc)
{
    return add_gemm<double>(b, transa, transb, m, n, k, alpha, d_A, lda, d_B, ldb, beta, d_C, ldc);
}

int gpublas_dev_batch_sgemm(gpublas_batch *b, char transa, char transb, int m, int n, int k,
                            float alpha, const float *d_A, int lda, const float *d_B, int ldb,
                            float beta, float *d_C, int ldc)
{
    return add_gemm<float>(b, transa, transb, m, n, k, alpha, d_A, lda, d_B, ldb, beta, d_C, ldc);
}

int gpublas_dev_batch_dgemv(gpublas_batch *b, char trans, int m, int n, double alpha,
                            const double *d_A, int lda, const double *d_x, int incx, double beta,
                            double *d_y, int incy)
{
    if (m <= 0 || n <= 0 || incx == 0 || incy == 0) return -1;
    GemvGroup &G = b->h_tab->gv;
    if (G.np >= kCapGemv) return 1;
    GGemv d;
    memset(&d, 0, sizeof d);
    d.A = d_A;
    d.x = d_x;
    d.y = d_y;
    d.lda = lda;
    d.incx = incx;
    d.incy = incy;
    d.m = m;
    d.n = n;
    d.trans = !is_n(trans);
    d.alpha = alpha;
    d.beta = beta;
    d.vec_ok = (((uintptr_t)d_A & 15) == 0) && (lda % 2 == 0);
    d.x_vec = (((uintptr_t)d_x & 15) == 0) && incx == 1;
    int units, len;
    if (!d.trans) {
        // N: 512-row blocks; column blocks (>= 64 columns) for about 512 units
        const int nrb = ceil_div(m, GV_ROWS);
        int ncb = ceil_div(512, nrb);
        const int maxcb = ceil_div(n, 64);
        if (ncb > maxcb) ncb = maxcb;
        if (ncb < 1) ncb = 1;
        d.span = ceil_div(n, ncb);
        d.split = ceil_div(n, d.span);
        units = nrb * d.split;
        len = m;
    } else {
        // T: 16-column units; row chunks (multiples of 512 rows) for about 256 units
        const int ncb = ceil_div(n, GV_TCOLS);
        int nrc = ceil_div(256, ncb);
        const int maxrc = ceil_div(m, 512);
        if (nrc > maxrc) nrc = maxrc;
        if (nrc < 1) nrc = 1;
        d.span = ceil_div(ceil_div(m, nrc), 512) * 512;
        d.split = ceil_div(m, d.span);
        units = ncb * d.split;
        len = n;
    }
    const int lenx = d.trans ? m : n;
    d.x0 = incx > 0 ? 0 : (long)(lenx - 1) * (long)(-incx);
    d.y0 = incy > 0 ? 0 : (long)(len - 1) * (long)(-incy);
    d.direct = d.split == 1;
    if (!d.direct) {
        const size_t need = (size_t)d.split * (size_t)len;
        if (b->ws_used + need > b->ws_cap) return G.np ? 1 : -1;
        d.part = b->d_ws + b->ws_used;
        b->ws_used += (need + 31) / 32 * 32;
    }
    d.unit_begin = G.units;
    d.fold_begin = G.fold_units;
    G.units += units;
    if (!d.direct) G.fold_units += ceil_div(len, GV_FOLD);
    G.p[G.np++] = d;
    return 0;
}

int gpublas_dev_batch_dtrsm(gpublas_batch *b, char side, char uplo, char transa, char diag, int m,
                            int n, double alpha, const double *d_A, int lda, double *d_B, int ldb)
{
    if (!is_l(side) || m <= 0 || m > TR_MAXM || n <= 0 || alpha == 0.0) return -1;
    TrsmGroup &G = b->h_tab->tr;
    if (G.np >= kCapTrsm) return 1;
    GTrsm d;
    memset(&d, 0, sizeof d);
    d.A = d_A;
    d.B = d_B;
    d.lda = lda;
    d.ldb = ldb;
    d.m = m;
    d.n = n;
    d.trans = !is_n(transa);
    d.fwd = (!is_u(uplo)) == !d.trans;  // op(A) lower
    d.unit = diag == 'U' || diag == 'u';
    d.alpha = alpha;
    d.unit_begin = G.units;
    G.units += ceil_div(n, 16);
    G.p[G.np++] = d;
    return 0;
}

int gpublas_dev_batch_launch(gpublas_batch *b)
{
    const unsigned mask = batch_mask(b->h_tab);
    b->launched = 1;
    int err = 0;
    PROF_BEGIN(b->s, "serve_batch");
    if (b->use_graph) {
        hipGraphExec_t e = graph_for(b, mask);
        if (!e) return (int)hipErrorUnknown;
        err = (int)hipGraphLaunch(e, b->s);
        b->graph_launches++;
    } else {
        enqueue(b, mask, false);
        err = (int)hipGetLastError();
        b->eager_launches++;
    }
    PROF_END(b->s);
    if (!err) err = (int)hipEventRecord(b->ev_done, b->s);
    return err;
}

int gpublas_dev_batch_done(gpublas_batch *b)
{
    if (!b->launched) return 1;
    return hipEventQuery(b->ev_done) == hipSuccess;
}

int gpublas_dev_batch_done_kind(gpublas_batch *b, int kind)
{
    if (!b->launched) return 1;
    if (b->use_graph || kind < 0 || kind > 3) return gpublas_dev_batch_done(b);
    if (!b->kind_used[kind]) return 1;
    return hipEventQuery(b->ev_kind[kind]) == hipSuccess;
}

int gpublas_dev_batch_wait(gpublas_batch *b)
{
    if (!b->launched) return 0;
    return (int)hipEventSynchronize(b->ev_done);
}

// Tile-shape table: measure each valid configuration on groups of 8
// problems per (sqrt(mn), k) bucket and keep the fastest.  `path`: load
// from it if it exists (unless force), else measure and write it.
int gpublas_dev_tune_gemm(const char *path, int force)
{
    if (path && !force) {
        FILE *f = fopen(path, "r");
        if (f) {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            tune_init_locked();
            char ty;
            int i, j, c, nread = 0;
            while (fscanf(f, " %c %d %d %d", &ty, &i, &j, &c) == 4) {
                if ((ty == 'd' || ty == 's') && i >= 0 && i < kB && j >= 0 && j < kB && c >= 0 &&
                    c < kCfgs) {
                    g_tune[ty == 'd' ? TY_D : TY_S][i][j] = (signed char)c;
                    ++nread;
                }
            }
            fclose(f);
            return nread > 0 ? 0 : -1;
        }
    }
    gpublas_batch *b = gpublas_dev_batch_create(0);
    if (!b) return -1;
    const int P = 8;
    const size_t maxe = (size_t)2048 * 2048;
    double *buf = nullptr;
    if (hipMalloc((void **)&buf, 3 * P * maxe * sizeof(double)) != hipSuccess) {
        gpublas_dev_batch_destroy(b);
        return -1;
    }
    (void)gpublas_dev_fill_uniform(b->s, 7, (int64_t)(3 * P * maxe), buf);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    signed char res[2][kB][kB];
    for (int ty = 0; ty < 2; ++ty)
        for (int i = 0; i < kB; ++i)
            for (int j = 0; j < kB; ++j) {
                const int mn = 32 << i, k = 32 << j;
                float best = 1e30f;
                int bestc = -1;
                for (int c = 0; c < kCfgs; ++c) {
                    if (!cfg_valid(ty, c)) continue;
                    float tot = 0;
                    for (int rep = 0; rep < 4; ++rep) {
                        batch_clear(b);
                        b->force_cfg = c;
                        for (int p = 0; p < P; ++p) {
                            double *base = buf + (size_t)3 * p * maxe;
                            if (ty == TY_D)
                                add_gemm<double>(b, 'N', 'N', mn, mn, k, 1.0, base, mn, base + maxe,
                                                 k, 0.0, base + 2 * maxe, mn);
                            else
                                add_gemm<float>(b, 'N', 'N', mn, mn, k, 1.0f, (const float *)base, mn,
                                                (const float *)(base + maxe), k, 0.0f,
                                                (float *)(base + 2 * maxe), mn);
                        }
                        const unsigned mask = batch_mask(b->h_tab);
                        (void)hipEventRecord(e0, b->s);
                        enqueue(b, mask, false);
                        (void)hipEventRecord(e1, b->s);
                        (void)hipStreamSynchronize(b->s);
                        float ms = 0;
                        (void)hipEventElapsedTime(&ms, e0, e1);
                        if (rep) tot += ms;
                    }
                    if (tot < best) {
                        best = tot;
                        bestc = c;
                    }
                }
                res[ty][i][j] = (signed char)bestc;
            }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(buf);
    gpublas_dev_batch_destroy(b);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        tune_init_locked();
        memcpy(g_tune, res, sizeof res);
    }
    if (path) {
        FILE *f = fopen(path, "w");
        if (f) {
            for (int ty = 0; ty < 2; ++ty)
                for (int i = 0; i < kB; ++i)
                    for (int j = 0; j < kB; ++j)
                        fprintf(f, "%c %d %d %d\n", ty ? 's' : 'd', i, j, res[ty][i][j]);
            fclose(f);
        }
    }
    return 0;
}

int gpublas_dev_tune_choice(int is_single, int m, int n, int k)
{
    return choose_cfg(is_single ? TY_S : TY_D, m, n, k);
}

void gpublas_dev_tune_reset(void)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    memset(g_tune, -1, sizeof g_tune);
    g_tune_init = true;
}

}  // extern "C"
