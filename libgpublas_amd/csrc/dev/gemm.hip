// gemm.hip -- LDS-tiled MFMA kernels for the real Level-3 routines on gfx950:
// DGEMM/SGEMM (all trans forms), DSYRK/SSYRK (one triangle), DSYMM/SSYMM.
//
// Replaces the vendor calls at blas_level3/gemm.cc:62-82 (cublasDgemm),
// syrk.cc:57-75 (cublasDsyrk) and symm.cc:57-78 (cublasDsymm).
//
// One tile routine (gemm_tile.hpp) serves all three routines, plus a
// specialised full-tile kernel for aligned rectangular operands:
//  * SYRK is the same product with B = A (transposed), restricted to tiles
//    on and below (above) the diagonal and masked to the triangle;
//  * SYMM reads the symmetric operand from its stored triangle: per k-tile
//    the loader picks the direct layout, the mirrored (transposed) layout,
//    or a per-element select for tiles that straddle the diagonal.
#include "gemm_tile.hpp"
#include <algorithm>
#include "gpublas_dev.h"

namespace {
template <typename T, int BM, int BN, int BK>
__global__ __launch_bounds__(256) void gemm_mfma_kernel(GemmArgs<T> g)
{
    constexpr int PAD = 128 / (int)sizeof(T);
    __shared__ T As[2][BK][BM + PAD];
    __shared__ T Bs[2][BK][BN + PAD];

    int tm, tn;
    if (g.tri == SYM_NONE) {
        tm = blockIdx.x % g.tiles_m;
        tn = blockIdx.x / g.tiles_m;
    } else {
        int hi, lo;
        tri_decode(blockIdx.x, hi, lo);
        if (g.tri == SYM_LOWER) { tm = hi; tn = lo; } else { tm = lo; tn = hi; }
    }
    if (tm < g.full_m && tn < g.full_n) return;  // interior tile: fast kernel's
    gemm_tile<T, BM, BN, BK>(g, tm, tn, As, Bs);
}

// A rectangular grid with both full and ragged tiles in one launch: full
// tiles take the fast tile routine (patterns PA / PB), ragged ones its
// guarded variant.  One launch instead of a fast grid followed by an edge
// grid, whose few tiles would otherwise make a round of their own.
template <typename T, int BM, int PA, int PB>
__global__ __launch_bounds__(256, 2) void gemm_mixed_kernel(GemmArgs<T> g)
{
    constexpr int PAD = 128 / (int)sizeof(T);
    __shared__ __attribute__((aligned(16))) T As[2][16][BM + PAD];
    __shared__ __attribute__((aligned(16))) T Bs[2][16][BM + PAD];
    const int tm = blockIdx.x % g.tiles_m, tn = blockIdx.x / g.tiles_m;
    if (tm < g.full_m && tn < g.full_n)
        gemm_fast_tile<T, BM, BM, 16, PA, PB, SYM_NONE>(g, tm, tn, As, Bs);
    else
        gemm_fast_tile<T, BM, BM, 16, PA, PB, SYM_NONE, true>(g, tm, tn, As, Bs);
}

// Fast kernel: the grid holds only tiles that are full in M and N, both
// operands are 16-B aligned with leading dimensions a multiple of the vector
// width, and their access patterns are compile-time constants (PA / PB).
// That leaves no per-element selects in the main loop and frees enough
// registers for two workgroups per CU, so one workgroup's global loads and
// barrier overlap the other's MFMAs.  Only a partial last k-tile takes the
// guarded element loads.
template <typename T, int BM, int BN, int BK, int PA, int PB, int TRI>
__global__ __launch_bounds__(256, 2) void gemm_fast_kernel(GemmArgs<T> g)
{
    constexpr int PAD = 128 / (int)sizeof(T);
    __shared__ __attribute__((aligned(16))) T As[2][BK][BM + PAD];
    __shared__ __attribute__((aligned(16))) T Bs[2][BK][BN + PAD];

    int tm, tn;
    if (TRI == SYM_NONE) {
        tm = blockIdx.x % g.tiles_m;
        tn = blockIdx.x / g.tiles_m;
    } else {
        int hi, lo;
        tri_decode(blockIdx.x, hi, lo);
        if (TRI == SYM_LOWER) { tm = hi; tn = lo; } else { tm = lo; tn = hi; }
    }
    gemm_fast_tile<T, BM, BN, BK, PA, PB, TRI>(g, tm, tn, As, Bs);
}

// The last, partly filled round of a 128 x 128 fast grid, run as 64 x 64
// quarter tiles: block b is quarter b % 4 of big tile base + b / 4 (same
// decoding as gemm_fast_kernel), so the round is four times shorter.
// Quarters wholly outside the stored triangle have nothing to do.  Same k
// order per element as the big tile: the bits do not change.
template <typename T, int PA, int PB, int TRI>
__global__ __launch_bounds__(256, 2) void gemm_fast_sub_kernel(GemmArgs<T> g, int base)
{
    constexpr int PAD = 128 / (int)sizeof(T);
    __shared__ __attribute__((aligned(16))) T As[2][16][64 + PAD];
    __shared__ __attribute__((aligned(16))) T Bs[2][16][64 + PAD];
    const int big = base + (int)(blockIdx.x >> 2), q = blockIdx.x & 3;
    int tm, tn;
    if (TRI == SYM_NONE) {
        tm = big % g.tiles_m;
        tn = big / g.tiles_m;
    } else {
        int hi, lo;
        tri_decode(big, hi, lo);
        if (TRI == SYM_LOWER) { tm = hi; tn = lo; } else { tm = lo; tn = hi; }
    }
    tm = 2 * tm + (q & 1);
    tn = 2 * tn + (q >> 1);
    if (TRI == SYM_LOWER && tn > tm) return;
    if (TRI == SYM_UPPER && tm > tn) return;
    gemm_fast_tile<T, 64, 64, 16, PA, PB, TRI>(g, tm, tn, As, Bs);
}

// Split-K for problems with too few tiles to fill the chip and a long k:
// block b is BM x BM tile b % tiles over k chunk b / tiles (kchunk deep, a
// multiple of 16), written unscaled to P[chunk] (m x n, leading dimension
// m); splitk_fold_kernel then forms alpha * (sum of chunks in chunk order) +
// beta * C.  The decomposition depends only on the shape, so the bits
// repeat from call to call.
template <typename T, int BM, int PA, int PB>
__global__ __launch_bounds__(256, 2) void gemm_fast_splitk_kernel(GemmArgs<T> g, int kchunk, T *P)
{
    constexpr int PAD = 128 / (int)sizeof(T);
    __shared__ __attribute__((aligned(16))) T As[2][16][BM + PAD];
    __shared__ __attribute__((aligned(16))) T Bs[2][16][BM + PAD];
    const int tiles = g.tiles_m * g.tiles_n;
    const int t = (int)(blockIdx.x % (unsigned)tiles), kc = (int)(blockIdx.x / (unsigned)tiles);
    const long koff = (long)kc * kchunk;
    GemmArgs<T> h = g;
    // element (o, k) of an operand is p[o + k*ld] (outer) or p[k + o*ld]
    h.a.p += pat_base(PA) == PAT_OUTER ? koff * g.a.ld : koff;
    h.b.p += pat_base(PB) == PAT_OUTER ? koff * g.b.ld : koff;
    h.k = kchunk;
    h.alpha = T(1);
    h.beta = T(0);
    h.C = P + (long)kc * g.m * g.n;
    h.ldc = g.m;
    gemm_fast_tile<T, BM, BM, 16, PA, PB, SYM_NONE>(h, t % g.tiles_m, t / g.tiles_m, As, Bs);
}

template <typename T>
__global__ __launch_bounds__(256) void splitk_fold_kernel(int m, int n, int ks, const T *P, T alpha,
                                                          T beta, T *C, long ldc)
{
    const long mn = (long)m * n;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < mn; e += (long)gridDim.x * 256) {
        T sum = P[e];
        for (int c = 1; c < ks; ++c) sum += P[(long)c * mn + e];
        const long i = e % m, j = e / m;
        T *cp = C + i + j * ldc;
        const T v = alpha * sum;
        *cp = beta == T(0) ? v : v + beta * *cp;
    }
}

// Scale (or zero) a triangle or rectangle of C: the alpha == 0 / k == 0
// paths of the netlib routines (beta == 0 stores exact zeros).
template <typename T>
__global__ void scale_kernel(int m, int n, T beta, T *C, long ldc, int tri)
{
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long total = (long)m * n;
    for (; idx < total; idx += (long)gridDim.x * blockDim.x) {
        long i = idx % m, j = idx / m;
        if (tri == SYM_LOWER && i < j) continue;
        if (tri == SYM_UPPER && i > j) continue;
        T *c = C + i + j * ldc;
        *c = beta == T(0) ? T(0) : beta * *c;
    }
}

template <typename T>
int launch_scale(hipStream_t s, int m, int n, T beta, T *C, long ldc, int tri)
{
    if (m <= 0 || n <= 0) return 0;
    long total = (long)m * n;
    int blocks = (int)((total + 255) / 256);
    if (blocks > 8192) blocks = 8192;
    PROF_BEGIN(s, "scale");
    hipLaunchKernelGGL(scale_kernel<T>, dim3(blocks), dim3(256), 0, s, m, n, beta, C, ldc, tri);
    PROF_END(s);
    return (int)hipGetLastError();
}

template <typename T> const char *kname(int big, int fast);
template <> const char *kname<double>(int big, int fast)
{
    if (fast) return big ? "dgemm_fast_128x128" : "dgemm_fast_64x64";
    return big ? "dgemm_edge_128x128" : "dgemm_edge_64x64";
}
template <> const char *kname<float>(int big, int fast)
{
    if (fast) return big ? "sgemm_fast_128x128" : "sgemm_fast_64x64";
    return big ? "sgemm_edge_128x128" : "sgemm_edge_64x64";
}

template <typename T, int BM>
void launch_fast(hipStream_t s, const GemmArgs<T> &g, unsigned blocks)
{
#define GB_FAST(PA_, PB_, TRI_)                                                                  \
    hipLaunchKernelGGL((gemm_fast_kernel<T, BM, BM, 16, PA_, PB_, TRI_>), dim3(blocks),         \
                       dim3(256), 0, s, g)
    const int pa = g.a.form, pb = g.b.form;
    const bool ua = !(g.a.vec_ok && g.b.vec_ok);  // element loads (rectangular grids only)
    if (g.tri == SYM_NONE && ua) {
        if (pa == PAT_OUTER && pb == PAT_KCONT) GB_FAST(PAT_OUTER_U, PAT_KCONT_U, SYM_NONE);
        else if (pa == PAT_OUTER) GB_FAST(PAT_OUTER_U, PAT_OUTER_U, SYM_NONE);
        else if (pb == PAT_KCONT) GB_FAST(PAT_KCONT_U, PAT_KCONT_U, SYM_NONE);
        else GB_FAST(PAT_KCONT_U, PAT_OUTER_U, SYM_NONE);
    } else if (g.tri == SYM_NONE) {
        if (pa == PAT_OUTER && pb == PAT_KCONT) GB_FAST(PAT_OUTER, PAT_KCONT, SYM_NONE);      // NN
        else if (pa == PAT_OUTER) GB_FAST(PAT_OUTER, PAT_OUTER, SYM_NONE);                    // NT
        else if (pb == PAT_KCONT) GB_FAST(PAT_KCONT, PAT_KCONT, SYM_NONE);                    // TN
        else GB_FAST(PAT_KCONT, PAT_OUTER, SYM_NONE);                                         // TT
    } else if (g.tri == SYM_LOWER) {
        if (pa == PAT_OUTER) GB_FAST(PAT_OUTER, PAT_OUTER, SYM_LOWER);  // A*A^T
        else GB_FAST(PAT_KCONT, PAT_KCONT, SYM_LOWER);                  // A^T*A
    } else {
        if (pa == PAT_OUTER) GB_FAST(PAT_OUTER, PAT_OUTER, SYM_UPPER);
        else GB_FAST(PAT_KCONT, PAT_KCONT, SYM_UPPER);
    }
#undef GB_FAST
}

template <typename T>
void launch_fast_sub(hipStream_t s, const GemmArgs<T> &g, int base, unsigned blocks)
{
#define GB_SUB(PA_, PB_, TRI_)                                                                   \
    hipLaunchKernelGGL((gemm_fast_sub_kernel<T, PA_, PB_, TRI_>), dim3(blocks), dim3(256), 0, s, \
                       g, base)
    const int pa = g.a.form, pb = g.b.form;
    const bool ua = !(g.a.vec_ok && g.b.vec_ok);
    if (g.tri == SYM_NONE && ua) {
        if (pa == PAT_OUTER && pb == PAT_KCONT) GB_SUB(PAT_OUTER_U, PAT_KCONT_U, SYM_NONE);
        else if (pa == PAT_OUTER) GB_SUB(PAT_OUTER_U, PAT_OUTER_U, SYM_NONE);
        else if (pb == PAT_KCONT) GB_SUB(PAT_KCONT_U, PAT_KCONT_U, SYM_NONE);
        else GB_SUB(PAT_KCONT_U, PAT_OUTER_U, SYM_NONE);
    } else if (g.tri == SYM_NONE) {
        if (pa == PAT_OUTER && pb == PAT_KCONT) GB_SUB(PAT_OUTER, PAT_KCONT, SYM_NONE);
        else if (pa == PAT_OUTER) GB_SUB(PAT_OUTER, PAT_OUTER, SYM_NONE);
        else if (pb == PAT_KCONT) GB_SUB(PAT_KCONT, PAT_KCONT, SYM_NONE);
        else GB_SUB(PAT_KCONT, PAT_OUTER, SYM_NONE);
    } else if (g.tri == SYM_LOWER) {
        if (pa == PAT_OUTER) GB_SUB(PAT_OUTER, PAT_OUTER, SYM_LOWER);
        else GB_SUB(PAT_KCONT, PAT_KCONT, SYM_LOWER);
    } else {
        if (pa == PAT_OUTER) GB_SUB(PAT_OUTER, PAT_OUTER, SYM_UPPER);
        else GB_SUB(PAT_KCONT, PAT_KCONT, SYM_UPPER);
    }
#undef GB_SUB
}

// Concurrent workgroup slots of the fast kernels: two per CU.
long fast_slots()
{
    static const long slots = [] {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0) {
            (void)hipGetLastError();
            cus = 256;
        }
        return 2L * cus;
    }();
    return slots;
}

// Split-K launch of a rectangular product whose BM x BM tiles are all full
// and number at most half the workgroup slots; k is cut into as many chunks
// as fill the slots (at most 8, each at least 256 deep, k divisible).
// Returns -1 when the shape does not qualify, so the caller takes the
// ordinary grid.
template <typename T, int BM>
int launch_splitk(hipStream_t s, const GemmArgs<T> &g, int m, int n, int k)
{
    const long tiles = (long)(m / BM) * (n / BM), slots = fast_slots();
    if (m % BM || n % BM || tiles == 0 || 2 * tiles > slots) return -1;
    int ks = (int)std::min<long>(slots / tiles, 8);
    while (ks > 1 && (k % (16 * ks) || k / ks < 256)) --ks;
    if (ks < 2) return -1;
    T *P = (T *)gpublas_dev_alloc((size_t)ks * m * n * sizeof(T));
    if (!P) return -1;
    GemmArgs<T> f = g;
    f.tiles_m = m / BM;
    f.tiles_n = n / BM;
    f.full_m = f.full_n = 0;
    const unsigned blocks = (unsigned)(tiles * ks);
#define GB_SK(PA_, PB_)                                                                          \
    hipLaunchKernelGGL((gemm_fast_splitk_kernel<T, BM, PA_, PB_>), dim3(blocks), dim3(256), 0, s, \
                       f, k / ks, P)
    const int pa = g.a.form, pb = g.b.form;
    const bool ua = !(g.a.vec_ok && g.b.vec_ok);
    PROF_BEGIN(s, sizeof(T) == 8 ? "dgemm_splitk" : "sgemm_splitk");
    if (ua) {
        if (pa == PAT_OUTER && pb == PAT_KCONT) GB_SK(PAT_OUTER_U, PAT_KCONT_U);
        else if (pa == PAT_OUTER) GB_SK(PAT_OUTER_U, PAT_OUTER_U);
        else if (pb == PAT_KCONT) GB_SK(PAT_KCONT_U, PAT_KCONT_U);
        else GB_SK(PAT_KCONT_U, PAT_OUTER_U);
    } else {
        if (pa == PAT_OUTER && pb == PAT_KCONT) GB_SK(PAT_OUTER, PAT_KCONT);
        else if (pa == PAT_OUTER) GB_SK(PAT_OUTER, PAT_OUTER);
        else if (pb == PAT_KCONT) GB_SK(PAT_KCONT, PAT_KCONT);
        else GB_SK(PAT_KCONT, PAT_OUTER);
    }
#undef GB_SK
    const long mn = (long)m * n;
    const unsigned fb = (unsigned)std::min<long>((mn + 255) / 256, 4096);
    hipLaunchKernelGGL(splitk_fold_kernel<T>, dim3(fb), dim3(256), 0, s, m, n, ks, (const T *)P,
                       g.alpha, g.beta, g.C, g.ldc);
    PROF_END(s);
    dev_release_after(P, s);
    return (int)hipGetLastError();
}

// Core launcher: C = alpha*op(A)*op(B) + beta*C on an (optionally
// triangular) tile grid.  k > 0, alpha != 0 assumed.  Tiles that are full in
// M and N go to the fast kernel (vector loads when both operands are
// aligned, element loads otherwise; a triangular grid needs aligned
// operands); the generic kernel takes the ragged edge (and everything, for
// symmetric operands).
template <typename T>
int launch_gemm(hipStream_t s, int m, int n, int k, T alpha, const Operand<T> &a,
                const Operand<T> &b, T beta, T *C, long ldc, int tri)
{
    GemmArgs<T> g;
    g.m = m; g.n = n; g.k = k;
    g.alpha = alpha; g.beta = beta;
    g.a = a; g.b = b;
    g.C = C; g.ldc = ldc;
    g.tri = tri;
    // Tile choice: 128x128 when that alone gives at least one tile per CU,
    // else 64x64 so that mid-size problems still fill the 256 CUs.
    const long t128 = (long)ceil_div(m, 128) * ceil_div(n, 128);
    const bool big = t128 >= 256;
    const int bm = big ? 128 : 64;
    // full tiles take the fast kernels: with vector loads when both operands
    // are aligned, else (rectangular grids) with element loads
    const bool fast_ok = a.sym == SYM_NONE && b.sym == SYM_NONE &&
                         ((a.vec_ok && b.vec_ok) || tri == SYM_NONE);
    if (fast_ok && !big && tri == SYM_NONE) {
        // few 64 x 64 tiles, long k: split k to fill the workgroup slots
        // (measured: 1024^3 32.8 -> 35.1 TFLOP/s, 512 x 512 x 4096 -> 37.4;
        // splitting 128 x 128 tiles instead lost to the larger fold)
        const int err = launch_splitk<T, 64>(s, g, m, n, k);
        if (err >= 0) return err;
    }
    const int full_m = fast_ok ? m / bm : 0, full_n = fast_ok ? n / bm : 0;
    const long fast_blocks = tri ? (long)full_m * (full_m + 1) / 2 : (long)full_m * full_n;
    if (tri == SYM_NONE && fast_blocks > 0 && (m % bm || n % bm)) {
        g.tiles_m = ceil_div(m, bm);
        g.tiles_n = ceil_div(n, bm);
        g.full_m = full_m;
        g.full_n = full_n;
        const unsigned blocks = (unsigned)((long)g.tiles_m * g.tiles_n);
        const bool ua = !(a.vec_ok && b.vec_ok);
#define GB_MIX(BM_, PA_, PB_)                                                                            hipLaunchKernelGGL((gemm_mixed_kernel<T, BM_, PA_, PB_>), dim3(blocks), dim3(256), 0, s, g)
#define GB_MIX4(BM_, O, K)                                                                               if (a.form == PAT_OUTER && b.form == PAT_KCONT) GB_MIX(BM_, O, K);                               else if (a.form == PAT_OUTER) GB_MIX(BM_, O, O);                                                 else if (b.form == PAT_KCONT) GB_MIX(BM_, K, K);                                                 else GB_MIX(BM_, K, O)
        PROF_BEGIN(s, kname<T>(big, 1));
        if (big) {
            if (ua) { GB_MIX4(128, PAT_OUTER_U, PAT_KCONT_U); }
            else { GB_MIX4(128, PAT_OUTER, PAT_KCONT); }
        } else {
            if (ua) { GB_MIX4(64, PAT_OUTER_U, PAT_KCONT_U); }
            else { GB_MIX4(64, PAT_OUTER, PAT_KCONT); }
        }
        PROF_END(s);
#undef GB_MIX4
#undef GB_MIX
        return (int)hipGetLastError();
    }
    if (fast_blocks > 0) {
        GemmArgs<T> f = g;
        f.tiles_m = full_m;
        f.tiles_n = full_n;
        f.full_m = f.full_n = 0;
        // A last round of 128 x 128 tiles that would leave most of the chip
        // idle runs as 64 x 64 quarters instead (a quarter of the time).
        const long slots = fast_slots(), tail = fast_blocks % slots;
        const long main_blocks =
            (big && fast_blocks > slots && 2 * tail < slots) ? fast_blocks - tail : fast_blocks;
        PROF_BEGIN(s, kname<T>(big, 1));
        if (big) launch_fast<T, 128>(s, f, (unsigned)main_blocks);
        else launch_fast<T, 64>(s, f, (unsigned)fast_blocks);
        if (main_blocks < fast_blocks)
            launch_fast_sub<T>(s, f, (int)main_blocks, (unsigned)(4 * (fast_blocks - main_blocks)));
        PROF_END(s);
    }
    g.tiles_m = ceil_div(m, bm);
    g.tiles_n = ceil_div(n, bm);
    g.full_m = full_m;
    g.full_n = full_n;
    const long blocks = tri ? (long)g.tiles_m * (g.tiles_m + 1) / 2 : (long)g.tiles_m * g.tiles_n;
    if (blocks > fast_blocks) {
        PROF_BEGIN(s, kname<T>(big, 0));
        if (big)
            hipLaunchKernelGGL((gemm_mfma_kernel<T, 128, 128, 16>), dim3((unsigned)blocks),
                               dim3(256), 0, s, g);
        else
            hipLaunchKernelGGL((gemm_mfma_kernel<T, 64, 64, 16>), dim3((unsigned)blocks),
                               dim3(256), 0, s, g);
        PROF_END(s);
    }
    return (int)hipGetLastError();
}

template <typename T>
int gemm_entry(void *stream, char transa, char transb, int m, int n, int k, T alpha, const T *A,
               int lda, const T *B, int ldb, T beta, T *C, int ldc)
{
    hipStream_t s = pick_stream(stream);
    if (m <= 0 || n <= 0) return 0;
    if (alpha == T(0) || k <= 0) {
        if (beta == T(1)) return 0;
        return launch_scale<T>(s, m, n, beta, C, ldc, SYM_NONE);
    }
    Operand<T> a = make_op<T>(A, lda, m, is_n(transa) ? PAT_OUTER : PAT_KCONT, SYM_NONE);
    Operand<T> b = make_op<T>(B, ldb, n, is_n(transb) ? PAT_KCONT : PAT_OUTER, SYM_NONE);
    return launch_gemm<T>(s, m, n, k, alpha, a, b, beta, C, ldc, SYM_NONE);
}

template <typename T>
int syrk_entry(void *stream, char uplo, char trans, int n, int k, T alpha, const T *A, int lda,
               T beta, T *C, int ldc)
{
    hipStream_t s = pick_stream(stream);
    const int tri = is_u(uplo) ? SYM_UPPER : SYM_LOWER;
    if (n <= 0) return 0;
    if (alpha == T(0) || k <= 0) {
        if (beta == T(1)) return 0;
        return launch_scale<T>(s, n, n, beta, C, ldc, tri);
    }
    // trans N: C = A*A^T  -> op(A) = A (n x k), op(B) = A^T
    // trans T: C = A^T*A  -> op(A) = A^T,       op(B) = A
    const bool nt = is_n(trans);
    Operand<T> a = make_op<T>(A, lda, n, nt ? PAT_OUTER : PAT_KCONT, SYM_NONE);
    Operand<T> b = make_op<T>(A, lda, n, nt ? PAT_OUTER : PAT_KCONT, SYM_NONE);
    return launch_gemm<T>(s, n, n, k, alpha, a, b, beta, C, ldc, tri);
}

// Full symmetric matrix S (n x n, leading dimension lds) from the stored
// triangle of A: one workgroup per 64 x 64 tile of S.  Tiles inside the
// stored triangle are copied, tiles in the other one are the transposed
// mirror tile (through LDS, so both sides stay coalesced), diagonal tiles
// pick per element.  Every thread issues its 16 loads before its stores.
template <typename T>
__global__ __launch_bounds__(256) void symmetrize_kernel(int n, const T *__restrict__ A, long lda,
                                                         int lower, T *__restrict__ S, long lds)
{
    __shared__ T t[64][65];
    const int I = blockIdx.x, J = blockIdx.y;
    const int i0 = I * 64, j0 = J * 64;
    const bool diag = I == J, direct = lower ? I > J : I < J;
    // source tile: (I, J) itself, or its mirror (J, I); element (a, b) of it
    // is A[(r0 + a) + (c0 + b) * lda]
    const int r0 = direct || diag ? i0 : j0, c0 = direct || diag ? j0 : i0;
    T v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int idx = threadIdx.x + 256 * u, a = idx & 63, b = idx >> 6;
        v[u] = (r0 + a < n && c0 + b < n) ? A[(r0 + a) + (long)(c0 + b) * lda] : T(0);
    }
    if (direct) {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int idx = threadIdx.x + 256 * u, a = idx & 63, b = idx >> 6;
            if (i0 + a < n && j0 + b < n) S[(i0 + a) + (long)(j0 + b) * lds] = v[u];
        }
        return;
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int idx = threadIdx.x + 256 * u;
        t[idx & 63][idx >> 6] = v[u];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int idx = threadIdx.x + 256 * u, a = idx & 63, b = idx >> 6;
        if (i0 + a >= n || j0 + b >= n) continue;
        // S(i0+a, j0+b): mirror tile -> A(j0+b, i0+a) = t[b][a]; diagonal
        // tile -> the stored one of (a, b) and (b, a)
        const bool stored = lower ? a >= b : a <= b;
        S[(i0 + a) + (long)(j0 + b) * lds] = (diag && stored) ? t[a][b] : t[b][a];
    }
}

// C = alpha*op(A)*op(B) where one factor is a full-tile triangle: tile
// (tm, tn) runs only over the k range where that factor is nonzero (the
// other products are exact zeros), so a triangular multiply costs half a
// GEMM.  KR: 1 k < (tm+1)*BM, 2 k >= tm*BM (left factor lower / upper),
// 3 k >= tn*BM, 4 k < (tn+1)*BM (right factor lower / upper).  Operand
// pointers are shifted to the range start as in the split-K kernel; tiles
// are dispatched longest range first.
template <typename T, int BM, int PA, int PB, int KR>
__global__ __launch_bounds__(256, 2) void gemm_fast_krange_kernel(GemmArgs<T> g)
{
    constexpr int PAD = 128 / (int)sizeof(T);
    __shared__ __attribute__((aligned(16))) T As[2][16][BM + PAD];
    __shared__ __attribute__((aligned(16))) T Bs[2][16][BM + PAD];
    // longest ranges first, so the last round holds the shortest tiles
    const int b = (int)blockIdx.x;
    int tm, tn;
    if (KR == 1 || KR == 2) {
        tm = b / g.tiles_n;
        if (KR == 1) tm = g.tiles_m - 1 - tm;
        tn = b % g.tiles_n;
    } else {
        tn = b / g.tiles_m;
        if (KR == 4) tn = g.tiles_n - 1 - tn;
        tm = b % g.tiles_m;
    }
    int kb = 0, ke = g.k;
    if (KR == 1) ke = min(g.k, (tm + 1) * BM);
    if (KR == 2) kb = tm * BM;
    if (KR == 3) kb = tn * BM;
    if (KR == 4) ke = min(g.k, (tn + 1) * BM);
    GemmArgs<T> h = g;
    h.a.p += PA == PAT_OUTER ? (long)kb * g.a.ld : kb;
    h.b.p += PB == PAT_OUTER ? (long)kb * g.b.ld : kb;
    h.k = ke - kb;
    gemm_fast_tile<T, BM, BM, 16, PA, PB, SYM_NONE>(h, tm, tn, As, Bs);
}

// Full n x n matrix of a triangle of A: the stored triangle copied, zeros
// elsewhere, ones on the diagonal when `unit` (the diagonal of A is then not
// read).  One workgroup per 64 x 64 tile; coalesced on both sides.
template <typename T>
__global__ __launch_bounds__(256) void triangle_kernel(int n, const T *__restrict__ A, long lda,
                                                       int lower, int unit, T *__restrict__ Tm,
                                                       long ldt)
{
    const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 64;
    T v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int idx = threadIdx.x + 256 * u, i = i0 + (idx & 63), j = j0 + (idx >> 6);
        const bool in = i < n && j < n && (lower ? i >= j : i <= j) && !(unit && i == j);
        v[u] = in ? A[i + (long)j * lda] : (unit && i == j ? T(1) : T(0));
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int idx = threadIdx.x + 256 * u, i = i0 + (idx & 63), j = j0 + (idx >> 6);
        if (i < n && j < n) Tm[i + (long)j * ldt] = v[u];
    }
}

template <typename T>
int symm_entry(void *stream, char side, char uplo, int m, int n, T alpha, const T *A, int lda,
               const T *B, int ldb, T beta, T *C, int ldc)
{
    hipStream_t s = pick_stream(stream);
    if (m <= 0 || n <= 0) return 0;
    if (alpha == T(0)) {
        if (beta == T(1)) return 0;
        return launch_scale<T>(s, m, n, beta, C, ldc, SYM_NONE);
    }
    const int sym = is_u(uplo) ? SYM_UPPER : SYM_LOWER;
    // From order 256 up, the stored triangle is first expanded into a full
    // symmetric matrix (one pass over na^2 elements) so the product runs on
    // the fast GEMM kernels; same products in the same k order, same bits.
    const int na = is_l(side) ? m : n;
    if (na >= 256) {
        const long lds = (na + 7) / 8 * 8;
        T *S = (T *)gpublas_dev_alloc((size_t)lds * na * sizeof(T));
        if (S) {
            const unsigned tiles = (unsigned)ceil_div(na, 64);
            PROF_BEGIN(s, "symmetrize");
            hipLaunchKernelGGL(symmetrize_kernel<T>, dim3(tiles, tiles), dim3(256), 0, s, na, A,
                               (long)lda, sym == SYM_LOWER ? 1 : 0, S, lds);
            PROF_END(s);
            int err = (int)hipGetLastError();
            if (!err) {
                if (is_l(side))
                    err = launch_gemm<T>(s, m, n, m, alpha,
                                         make_op<T>(S, lds, m, PAT_OUTER, SYM_NONE),
                                         make_op<T>(B, ldb, n, PAT_KCONT, SYM_NONE), beta, C, ldc,
                                         SYM_NONE);
                else
                    err = launch_gemm<T>(s, m, n, n, alpha,
                                         make_op<T>(B, ldb, m, PAT_OUTER, SYM_NONE),
                                         make_op<T>(S, lds, n, PAT_KCONT, SYM_NONE), beta, C, ldc,
                                         SYM_NONE);
            }
            dev_release_after(S, s);
            return err;
        }
    }
    if (is_l(side)) {
        // C = alpha*A*B + beta*C, A symmetric m x m: k = m
        Operand<T> a = make_op<T>(A, lda, m, PAT_OUTER, sym);
        Operand<T> b = make_op<T>(B, ldb, n, PAT_KCONT, SYM_NONE);
        return launch_gemm<T>(s, m, n, m, alpha, a, b, beta, C, ldc, SYM_NONE);
    }
    // C = alpha*B*A + beta*C, A symmetric n x n: k = n
    Operand<T> a = make_op<T>(B, ldb, m, PAT_OUTER, SYM_NONE);
    Operand<T> b = make_op<T>(A, lda, n, PAT_KCONT, sym);
    return launch_gemm<T>(s, m, n, n, alpha, a, b, beta, C, ldc, SYM_NONE);
}

// Triangular-k product (see gemm_fast_krange_kernel); -1 when the shape or
// the operands do not allow the fast tiles (the caller then runs the
// ordinary full-k GEMM).  beta = 0.
template <typename T, int BM>
int launch_krange(hipStream_t s, int m, int n, int k, T alpha, const Operand<T> &a,
                  const Operand<T> &b, T *C, long ldc, int kr)
{
    if (m % BM || n % BM || !a.vec_ok || !b.vec_ok) return -1;
    GemmArgs<T> g;
    g.m = m; g.n = n; g.k = k;
    g.alpha = alpha; g.beta = T(0);
    g.a = a; g.b = b;
    g.C = C; g.ldc = ldc;
    g.tri = SYM_NONE;
    g.tiles_m = m / BM;
    g.tiles_n = n / BM;
    g.full_m = g.full_n = 0;
    const unsigned blocks = (unsigned)(g.tiles_m * g.tiles_n);
#define GB_KR(PA_, PB_)                                                                          \
    do {                                                                                         \
        if (kr == 1) hipLaunchKernelGGL((gemm_fast_krange_kernel<T, BM, PA_, PB_, 1>), dim3(blocks), dim3(256), 0, s, g); \
        else if (kr == 2) hipLaunchKernelGGL((gemm_fast_krange_kernel<T, BM, PA_, PB_, 2>), dim3(blocks), dim3(256), 0, s, g); \
        else if (kr == 3) hipLaunchKernelGGL((gemm_fast_krange_kernel<T, BM, PA_, PB_, 3>), dim3(blocks), dim3(256), 0, s, g); \
        else hipLaunchKernelGGL((gemm_fast_krange_kernel<T, BM, PA_, PB_, 4>), dim3(blocks), dim3(256), 0, s, g); \
    } while (0)
    const int pa = a.form, pb = b.form;
    PROF_BEGIN(s, sizeof(T) == 8 ? "dgemm_krange" : "sgemm_krange");
    if (pa == PAT_OUTER && pb == PAT_KCONT) GB_KR(PAT_OUTER, PAT_KCONT);
    else if (pa == PAT_OUTER) GB_KR(PAT_OUTER, PAT_OUTER);
    else if (pb == PAT_KCONT) GB_KR(PAT_KCONT, PAT_KCONT);
    else GB_KR(PAT_KCONT, PAT_OUTER);
#undef GB_KR
    PROF_END(s);
    return (int)hipGetLastError();
}

// C = alpha*op(A)*op(B)^T + alpha*op(B)*op(A)^T + beta*C on the uplo triangle
// (op(X) = X, n x k, for trans N; X^T for trans T): two triangular-grid
// products, the second accumulating onto the first.
template <typename T>
int syr2k_entry(void *stream, char uplo, char trans, int n, int k, T alpha, const T *A, int lda,
                const T *B, int ldb, T beta, T *C, int ldc)
{
    hipStream_t s = pick_stream(stream);
    const int tri = is_u(uplo) ? SYM_UPPER : SYM_LOWER;
    if (n <= 0) return 0;
    if (alpha == T(0) || k <= 0) {
        if (beta == T(1)) return 0;
        return launch_scale<T>(s, n, n, beta, C, ldc, tri);
    }
    const bool nt = is_n(trans);
    Operand<T> a = make_op<T>(A, lda, n, nt ? PAT_OUTER : PAT_KCONT, SYM_NONE);
    Operand<T> b = make_op<T>(B, ldb, n, nt ? PAT_OUTER : PAT_KCONT, SYM_NONE);
    int err = launch_gemm<T>(s, n, n, k, alpha, a, b, beta, C, ldc, tri);
    if (!err) err = launch_gemm<T>(s, n, n, k, alpha, b, a, T(1), C, ldc, tri);
    return err;
}

// B := alpha*op(A)*B (side L) or alpha*B*op(A) (side R), A triangular: the
// triangle is expanded into a full matrix (zeros outside it, ones on a unit
// diagonal), the product runs on the GEMM tiles into a workspace -- each tile
// only over the k range where the triangle is nonzero when the tiles are
// full -- and the workspace is copied over B.
template <typename T>
int trmm_entry(void *stream, char side, char uplo, char transa, char diag, int m, int n, T alpha,
               const T *A, int lda, T *B, int ldb)
{
    hipStream_t s = pick_stream(stream);
    if (m <= 0 || n <= 0) return 0;
    if (alpha == T(0)) return launch_scale<T>(s, m, n, T(0), B, ldb, SYM_NONE);
    const bool left = is_l(side);
    const int na = left ? m : n;
    const long ldt = (na + 7) / 8 * 8, ldw = (m + 7) / 8 * 8;
    T *Tm = (T *)gpublas_dev_alloc((size_t)ldt * na * sizeof(T));
    T *W = (T *)gpublas_dev_alloc((size_t)ldw * n * sizeof(T));
    if (!Tm || !W) {
        gpublas_dev_release(Tm);
        gpublas_dev_release(W);
        return (int)hipErrorOutOfMemory;
    }
    const unsigned tiles = (unsigned)ceil_div(na, 64);
    PROF_BEGIN(s, "triangle");
    hipLaunchKernelGGL(triangle_kernel<T>, dim3(tiles, tiles), dim3(256), 0, s, na, A, (long)lda,
                       is_u(uplo) ? 0 : 1, (diag == 'U' || diag == 'u') ? 1 : 0, Tm, ldt);
    PROF_END(s);
    int err = (int)hipGetLastError();
    const bool tn = is_n(transa);
    // op(A) lower: the transposed upper triangle, or the lower one as stored
    const bool oplower = is_u(uplo) ? !tn : tn;
    if (!err) {
        const Operand<T> a = left ? make_op<T>(Tm, ldt, m, tn ? PAT_OUTER : PAT_KCONT, SYM_NONE)
                                  : make_op<T>(B, ldb, m, PAT_OUTER, SYM_NONE);
        const Operand<T> b = left ? make_op<T>(B, ldb, n, PAT_KCONT, SYM_NONE)
                                  : make_op<T>(Tm, ldt, n, tn ? PAT_KCONT : PAT_OUTER, SYM_NONE);
        const int k = left ? m : n;
        const int kr = left ? (oplower ? 1 : 2) : (oplower ? 3 : 4);
        const long t128 = (long)(m / 128) * (n / 128);
        err = t128 >= 256 ? launch_krange<T, 128>(s, m, n, k, alpha, a, b, W, ldw, kr)
                          : launch_krange<T, 64>(s, m, n, k, alpha, a, b, W, ldw, kr);
        if (err < 0) err = launch_gemm<T>(s, m, n, k, alpha, a, b, T(0), W, ldw, SYM_NONE);
    }
    if (!err)
        err = (int)hipMemcpy2DAsync(B, (size_t)ldb * sizeof(T), W, (size_t)ldw * sizeof(T),
                                    (size_t)m * sizeof(T), (size_t)n, hipMemcpyDeviceToDevice, s);
    dev_release_after(Tm, s);
    dev_release_after(W, s);
    return err;
}

}  // namespace

// Shared with trsm.hip: plain GEMM on device pointers.
int gpublas_gemm_d(hipStream_t s, char ta, char tb, int m, int n, int k, double alpha,
                   const double *A, int lda, const double *B, int ldb, double beta, double *C,
                   int ldc)
{
    return gemm_entry<double>(s, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
int gpublas_gemm_s(hipStream_t s, char ta, char tb, int m, int n, int k, float alpha,
                   const float *A, int lda, const float *B, int ldb, float beta, float *C, int ldc)
{
    return gemm_entry<float>(s, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

// Shared with cplx.hip: the uplo triangle of alpha op(A) op(B) + beta C on
// the triangular tile grid (op(A) n x k, op(B) k x n); the rest of C is not
// touched.
template <typename T>
int gemm_tri(hipStream_t s, char uplo, char ta, char tb, int n, int k, T alpha, const T *A,
             int lda, const T *B, int ldb, T beta, T *C, int ldc)
{
    const int tri = is_u(uplo) ? SYM_UPPER : SYM_LOWER;
    if (n <= 0) return 0;
    if (alpha == T(0) || k <= 0) {
        if (beta == T(1)) return 0;
        return launch_scale<T>(s, n, n, beta, C, ldc, tri);
    }
    Operand<T> a = make_op<T>(A, lda, n, is_n(ta) ? PAT_OUTER : PAT_KCONT, SYM_NONE);
    Operand<T> b = make_op<T>(B, ldb, n, is_n(tb) ? PAT_KCONT : PAT_OUTER, SYM_NONE);
    return launch_gemm<T>(s, n, n, k, alpha, a, b, beta, C, ldc, tri);
}
int gpublas_gemm_tri_d(hipStream_t s, char uplo, char ta, char tb, int n, int k, double alpha,
                       const double *A, int lda, const double *B, int ldb, double beta, double *C,
                       int ldc)
{
    return gemm_tri<double>(s, uplo, ta, tb, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
int gpublas_gemm_tri_s(hipStream_t s, char uplo, char ta, char tb, int n, int k, float alpha,
                       const float *A, int lda, const float *B, int ldb, float beta, float *C,
                       int ldc)
{
    return gemm_tri<float>(s, uplo, ta, tb, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

extern "C" {

int gpublas_dev_dgemm(void *stream, char transa, char transb, int m, int n, int k, double alpha,
                      const double *d_A, int lda, const double *d_B, int ldb, double beta,
                      double *d_C, int ldc)
{
    return gemm_entry<double>(stream, transa, transb, m, n, k, alpha, d_A, lda, d_B, ldb, beta,
                              d_C, ldc);
}

int gpublas_dev_sgemm(void *stream, char transa, char transb, int m, int n, int k, float alpha,
                      const float *d_A, int lda, const float *d_B, int ldb, float beta,
                      float *d_C, int ldc)
{
    return gemm_entry<float>(stream, transa, transb, m, n, k, alpha, d_A, lda, d_B, ldb, beta,
                             d_C, ldc);
}

int gpublas_dev_dsyrk(void *stream, char uplo, char trans, int n, int k, double alpha,
                      const double *d_A, int lda, double beta, double *d_C, int ldc)
{
    return syrk_entry<double>(stream, uplo, trans, n, k, alpha, d_A, lda, beta, d_C, ldc);
}

int gpublas_dev_ssyrk(void *stream, char uplo, char trans, int n, int k, float alpha,
                      const float *d_A, int lda, float beta, float *d_C, int ldc)
{
    return syrk_entry<float>(stream, uplo, trans, n, k, alpha, d_A, lda, beta, d_C, ldc);
}

int gpublas_dev_dsyr2k(void *stream, char uplo, char trans, int n, int k, double alpha,
                       const double *d_A, int lda, const double *d_B, int ldb, double beta,
                       double *d_C, int ldc)
{
    return syr2k_entry<double>(stream, uplo, trans, n, k, alpha, d_A, lda, d_B, ldb, beta, d_C, ldc);
}

int gpublas_dev_ssyr2k(void *stream, char uplo, char trans, int n, int k, float alpha,
                       const float *d_A, int lda, const float *d_B, int ldb, float beta,
                       float *d_C, int ldc)
{
    return syr2k_entry<float>(stream, uplo, trans, n, k, alpha, d_A, lda, d_B, ldb, beta, d_C, ldc);
}

int gpublas_dev_dtrmm(void *stream, char side, char uplo, char transa, char diag, int m, int n,
                      double alpha, const double *d_A, int lda, double *d_B, int ldb)
{
    return trmm_entry<double>(stream, side, uplo, transa, diag, m, n, alpha, d_A, lda, d_B, ldb);
}

int gpublas_dev_strmm(void *stream, char side, char uplo, char transa, char diag, int m, int n,
                      float alpha, const float *d_A, int lda, float *d_B, int ldb)
{
    return trmm_entry<float>(stream, side, uplo, transa, diag, m, n, alpha, d_A, lda, d_B, ldb);
}

int gpublas_dev_dsymm(void *stream, char side, char uplo, int m, int n, double alpha,
                      const double *d_A, int lda, const double *d_B, int ldb, double beta,
                      double *d_C, int ldc)
{
    return symm_entry<double>(stream, side, uplo, m, n, alpha, d_A, lda, d_B, ldb, beta, d_C, ldc);
}

int gpublas_dev_ssymm(void *stream, char side, char uplo, int m, int n, float alpha,
                      const float *d_A, int lda, const float *d_B, int ldb, float beta,
                      float *d_C, int ldc)
{
    return symm_entry<float>(stream, side, uplo, m, n, alpha, d_A, lda, d_B, ldb, beta, d_C, ldc);
}

}  // extern "C"
