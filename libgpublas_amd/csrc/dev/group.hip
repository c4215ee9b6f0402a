// group.hip -- grouped persistent launches for the serving loop: one batch
// slot holds up to kCap problems of each kind (DGEMM, SGEMM, DGEMV, DTRSM)
// of arbitrary, uneven sizes and runs them all with one replay of a
// captured hipGraph (or the same kernels launched eagerly).
//
// No counterpart in the reference: blas2cuda issues one cuBLAS call per
// intercepted BLAS call (blas_level3/gemm.cc:62-82, trsm.cc:52-72) on one
// global handle.  Here the host engine collects calls from many threads and
// hands them over in batches.
//
// Layout of one batch in HBM: a BatchTable (problem descriptors and
// per-group prefix sums of work units) copied from pinned host memory by the
// first graph node, a GEMV partial-sum workspace, and the operands
// themselves (device-resident tracked objects, addressed in place).
//
// Kernels (all persistent: a fixed grid loops over the batch's work units,
// so one captured graph serves every batch of the slot):
//  * group_gemm_kernel<T, BM, BK>: every tile of every GEMM problem that the
//    tile-shape table assigns to BM x BM; tile t -> problem by binary search
//    over the prefix sums; the tile itself is the per-call generic routine
//    (gemm_tile.hpp), so a product's bits do not depend on the batch;
//  * group_gemv_part_kernel / group_gemv_fold_kernel: y = alpha*op(A)*x +
//    beta*y as column-block (N) or row-chunk (T) partial sums followed by a
//    fixed-order fold; the split depends only on (m, n);
//  * group_trsm_kernel: left-side triangular solves (m <= 1024), one
//    workgroup per 16-column panel of B held in LDS, right-looking over
//    16-row blocks (in-register substitution of the diagonal block, MFMA
//    updates of the blocks below);
//  * publish_kernel: stores the batch sequence number into a pinned host
//    word that the host polls.
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

constexpr int kCap = 64;  // problems of each kind per batch
enum { CFG_32 = 0, CFG_64 = 1, CFG_128 = 2, kCfgs = 3 };
enum { TY_D = 0, TY_S = 1 };

template <typename T>
struct GGemm {
    GemmArgs<T> g;
    int tile_begin;
};

struct GGemv {
    const double *A;
    const double *x;
    double *y;
    double *part;  // split x len partial sums
    long lda;
    int incx, incy;
    int m, n, trans;
    int split;  // N: column blocks; T: row chunks
    int span;   // N: columns per block; T: rows per chunk
    int vec_ok;  // A 16-B aligned, lda even
    int x_vec;   // x 16-B aligned, incx == 1
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
    int unit_begin;  // waves
};

struct BatchTable {
    uint64_t seq;
    int gemm_n[2][kCfgs], gemm_first[2][kCfgs], gemm_tiles[2][kCfgs];
    int ngemv, gemv_units, fold_units;
    int ntrsm, trsm_units;
    GGemm<double> dg[kCap];
    GGemm<float> sg[kCap];
    GGemv gv[kCap];
    GTrsm tr[kCap];
};

template <typename T> __device__ __forceinline__ const GGemm<T> *gemm_list(const BatchTable *t);
template <> __device__ __forceinline__ const GGemm<double> *gemm_list<double>(const BatchTable *t)
{
    return t->dg;
}
template <> __device__ __forceinline__ const GGemm<float> *gemm_list<float>(const BatchTable *t)
{
    return t->sg;
}

// ---- grouped GEMM -----------------------------------------------------------
template <typename T, int BM, int BK, int CFG>
__global__ __launch_bounds__(256) void group_gemm_kernel(const BatchTable *tb)
{
    constexpr int PAD = 128 / (int)sizeof(T);
    __shared__ T As[2][BK][BM + PAD];
    __shared__ T Bs[2][BK][BM + PAD];
    constexpr int ty = sizeof(T) == 8 ? TY_D : TY_S;
    const int np = min(tb->gemm_n[ty][CFG], kCap);
    const int total = tb->gemm_tiles[ty][CFG];
    const GGemm<T> *P = gemm_list<T>(tb) + tb->gemm_first[ty][CFG];
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        int lo = 0, hi = np - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (P[mid].tile_begin <= t) lo = mid;
            else hi = mid - 1;
        }
        const GemmArgs<T> g = P[lo].g;
        const int tl = t - P[lo].tile_begin;
        gemm_tile<T, BM, BM, BK>(g, tl % g.tiles_m, tl / g.tiles_m, As, Bs);
    }
}

// ---- grouped GEMV -----------------------------------------------------------
constexpr int GV_ROWS = 512;    // N: rows per unit (2 per thread)
constexpr int GV_TCOLS = 16;    // T: columns per unit (4 per wave)
constexpr int GV_FOLD = 256;    // fold: outputs per unit

__device__ __forceinline__ long vidx(int i, int inc, int len)
{
    return inc > 0 ? (long)i * inc : (long)(len - 1 - i) * (long)(-inc);
}

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

// y-partial for rows [rb*512, +512) over columns [cb*span, +span)
__device__ void gemv_n_unit(const GGemv &d, int local)
{
    typedef double d2 __attribute__((ext_vector_type(2)));
    const int nrb = (d.m + GV_ROWS - 1) / GV_ROWS;
    const int rb = local % nrb, cb = local / nrb;
    const int r0 = rb * GV_ROWS + 2 * (int)threadIdx.x;
    const int c0 = cb * d.span, c1 = min(d.n, c0 + d.span);
    double acc0 = 0.0, acc1 = 0.0;
    if (r0 < d.m) {
        const bool pair = r0 + 1 < d.m;
        if (pair && d.vec_ok) {
            int j = c0;
            for (; j + 8 <= c1; j += 8) {
                d2 a[8];
                double xv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    a[u] = __builtin_nontemporal_load(
                        (const d2 *)(d.A + r0 + (long)(j + u) * d.lda));
                    xv[u] = d.x[vidx(j + u, d.incx, d.n)];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    acc0 = fma(a[u].x, xv[u], acc0);
                    acc1 = fma(a[u].y, xv[u], acc1);
                }
            }
            for (; j < c1; ++j) {
                const d2 a = *(const d2 *)(d.A + r0 + (long)j * d.lda);
                const double xv = d.x[vidx(j, d.incx, d.n)];
                acc0 = fma(a.x, xv, acc0);
                acc1 = fma(a.y, xv, acc1);
            }
        } else {
            for (int j = c0; j < c1; ++j) {
                const double xv = d.x[vidx(j, d.incx, d.n)];
                const double *a = d.A + r0 + (long)j * d.lda;
                acc0 = fma(a[0], xv, acc0);
                if (pair) acc1 = fma(a[1], xv, acc1);
            }
        }
        double *p = d.part + (long)cb * d.m + r0;
        p[0] = acc0;
        if (pair) p[1] = acc1;
    }
}

// partial dot products of 16 columns over rows [rc*span, +span)
__device__ void gemv_t_unit(const GGemv &d, int local)
{
    typedef double d2 __attribute__((ext_vector_type(2)));
    const int ncb = (d.n + GV_TCOLS - 1) / GV_TCOLS;
    const int cb = local % ncb, rc = local / ncb;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j0 = cb * GV_TCOLS + wave * 4;
    const int r0 = rc * d.span, r1 = min(d.m, r0 + d.span);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const bool vec = d.x_vec;
    for (int i = r0 + 2 * lane; i < r1; i += 128) {
        const bool pair = i + 1 < r1;
        double x0, x1 = 0.0;
        if (vec && pair) {
            const d2 xv = *(const d2 *)(d.x + i);
            x0 = xv.x;
            x1 = xv.y;
        } else {
            x0 = d.x[vidx(i, d.incx, d.m)];
            if (pair) x1 = d.x[vidx(i + 1, d.incx, d.m)];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int j = j0 + c;
            if (j >= d.n) continue;
            const double *a = d.A + i + (long)j * d.lda;
            double a0, a1 = 0.0;
            if (d.vec_ok && pair) {
                const d2 av = __builtin_nontemporal_load((const d2 *)a);
                a0 = av.x;
                a1 = av.y;
            } else {
                a0 = a[0];
                if (pair) a1 = a[1];
            }
            acc[c] = fma(a0, x0, acc[c]);
            if (pair) acc[c] = fma(a1, x1, acc[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        double v = acc[c];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0 && j0 + c < d.n) d.part[(long)rc * d.n + j0 + c] = v;
    }
}

__global__ __launch_bounds__(256) void group_gemv_part_kernel(const BatchTable *tb)
{
    const int np = min(tb->ngemv, kCap);
    const int total = tb->gemv_units;
    for (int u = blockIdx.x; u < total; u += gridDim.x) {
        const int p = find_unit(&tb->gv[0].unit_begin, sizeof(GGemv), np, u);
        const GGemv d = tb->gv[p];
        if (d.trans) gemv_t_unit(d, u - d.unit_begin);
        else gemv_n_unit(d, u - d.unit_begin);
    }
}

__global__ __launch_bounds__(256) void group_gemv_fold_kernel(const BatchTable *tb)
{
    const int np = min(tb->ngemv, kCap);
    const int total = tb->fold_units;
    for (int u = blockIdx.x; u < total; u += gridDim.x) {
        const int p = find_unit(&tb->gv[0].fold_begin, sizeof(GGemv), np, u);
        const GGemv d = tb->gv[p];
        const int len = d.trans ? d.n : d.m;
        const int i = (u - d.fold_begin) * GV_FOLD + (int)threadIdx.x;
        if (i >= len) continue;
        double s = 0.0;
        for (int q = 0; q < d.split; ++q) s += d.part[(long)q * len + i];
        double *y = d.y + vidx(i, d.incy, len);
        const double v = d.alpha * s;
        *y = d.beta == 0.0 ? v : v + d.beta * *y;
    }
}

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

__global__ __launch_bounds__(1024) void group_trsm_kernel(const BatchTable *tb)
{
    __shared__ double Bs[TR_MAXM][17];
    __shared__ double Ds[16][17];
    const int np = min(tb->ntrsm, kCap);
    const int total = tb->trsm_units;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int q0 = lane >> 4, cl = lane & 15;
    for (int u = blockIdx.x; u < total; u += gridDim.x) {
        const int p = find_unit(&tb->tr[0].unit_begin, sizeof(GTrsm), np, u);
        const GTrsm d = tb->tr[p];
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

__global__ void publish_kernel(const BatchTable *tb, uint64_t *flag)
{
    if (threadIdx.x == 0) {
        __threadfence_system();
        __hip_atomic_store(flag, tb->seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

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
// kernel's occupancy, capped by the work when launched eagerly
constexpr int kGridGemm[kCfgs] = {1024, 1024, 512};
constexpr int kGridGemv = 1024, kGridFold = 512, kGridTrsm = 256;  // TRSM: 1024-thread groups

}  // namespace

// ---- batch slot -------------------------------------------------------------
struct gpublas_batch {
    int use_graph;
    int force_cfg;
    hipStream_t s;
    hipStream_t br[3];
    hipEvent_t ev_fork, ev_join[3];
    BatchTable *h_tab;
    BatchTable *d_tab;
    double *d_ws;
    size_t ws_cap, ws_used;  // doubles
    uint64_t *h_flag;
    uint64_t seq;
    int launched;
    // staging, filled by add_*, laid out into h_tab at launch
    GemmArgs<double> dq[kCfgs][kCap];
    GemmArgs<float> sq[kCfgs][kCap];
    int nd[kCfgs], ns[kCfgs], nd_all, ns_all;
    hipGraphExec_t exec[256];
    long graph_launches, eager_launches;
};

namespace {

constexpr size_t kWsDoubles = (size_t)8 << 20;  // 64 MiB of GEMV partial sums

void batch_clear(gpublas_batch *b)
{
    for (int c = 0; c < kCfgs; ++c) b->nd[c] = b->ns[c] = 0;
    b->nd_all = b->ns_all = 0;
    b->h_tab->ngemv = b->h_tab->gemv_units = b->h_tab->fold_units = 0;
    b->h_tab->ntrsm = b->h_tab->trsm_units = 0;
    b->ws_used = 0;
}

template <typename T>
int add_gemm(gpublas_batch *b, char ta, char tb, int m, int n, int k, T alpha, const T *A, int lda,
             const T *B, int ldb, T beta, T *C, int ldc)
{
    constexpr int ty = sizeof(T) == 8 ? TY_D : TY_S;
    if (m <= 0 || n <= 0 || k <= 0 || alpha == T(0)) return -1;
    int &count = ty == TY_D ? b->nd_all : b->ns_all;
    if (count >= kCap) return 1;
    int cfg = b->force_cfg >= 0 && cfg_valid(ty, b->force_cfg) ? b->force_cfg
                                                               : choose_cfg(ty, m, n, k);
    GemmArgs<T> g;
    memset(&g, 0, sizeof g);
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
    g.full_m = g.full_n = 0;
    if constexpr (ty == TY_D) b->dq[cfg][b->nd[cfg]++] = g;
    else b->sq[cfg][b->ns[cfg]++] = g;
    ++count;
    return 0;
}

void lay_out(gpublas_batch *b, unsigned *mask)
{
    BatchTable *t = b->h_tab;
    int off = 0;
    for (int c = 0; c < kCfgs; ++c) {
        t->gemm_first[TY_D][c] = off;
        t->gemm_n[TY_D][c] = b->nd[c];
        int tiles = 0;
        for (int i = 0; i < b->nd[c]; ++i) {
            t->dg[off + i].g = b->dq[c][i];
            t->dg[off + i].tile_begin = tiles;
            tiles += b->dq[c][i].tiles_m * b->dq[c][i].tiles_n;
        }
        t->gemm_tiles[TY_D][c] = tiles;
        if (tiles) *mask |= 1u << c;
        off += b->nd[c];
    }
    off = 0;
    for (int c = 0; c < kCfgs; ++c) {
        t->gemm_first[TY_S][c] = off;
        t->gemm_n[TY_S][c] = b->ns[c];
        int tiles = 0;
        for (int i = 0; i < b->ns[c]; ++i) {
            t->sg[off + i].g = b->sq[c][i];
            t->sg[off + i].tile_begin = tiles;
            tiles += b->sq[c][i].tiles_m * b->sq[c][i].tiles_n;
        }
        t->gemm_tiles[TY_S][c] = tiles;
        if (tiles) *mask |= 1u << (3 + c);
        off += b->ns[c];
    }
    if (t->gemv_units) *mask |= 1u << 6;
    if (t->trsm_units) *mask |= 1u << 7;
}

template <typename T, int BM, int CFG>
void launch_gemm_group(hipStream_t s, const BatchTable *d_tab, int grid)
{
    hipLaunchKernelGGL((group_gemm_kernel<T, BM, 16, CFG>), dim3(grid), dim3(256), 0, s, d_tab);
}

// Enqueue the batch's kernels.  In graph mode the grids are the fixed
// persistent sizes (the table's counts are read on the device); eagerly
// they are capped by the work.
void enqueue(gpublas_batch *b, unsigned mask, bool capture)
{
    const BatchTable *t = b->h_tab;
    const BatchTable *d = b->d_tab;
    auto grid = [&](int fixed, int work) { return capture ? fixed : (work < fixed ? work : fixed); };
    hipStream_t s0 = b->s, sg = b->s, sv = b->s, st = b->s;
    (void)hipMemcpyAsync(b->d_tab, b->h_tab, sizeof(BatchTable), hipMemcpyHostToDevice, s0);
    if (capture) {
        // three branches: GEMM groups, GEMV, TRSM
        (void)hipEventRecord(b->ev_fork, s0);
        sg = b->br[0];
        sv = b->br[1];
        st = b->br[2];
        (void)hipStreamWaitEvent(sg, b->ev_fork, 0);
        (void)hipStreamWaitEvent(sv, b->ev_fork, 0);
        (void)hipStreamWaitEvent(st, b->ev_fork, 0);
    }
    if (mask & 1u) launch_gemm_group<double, 32, CFG_32>(sg, d, grid(kGridGemm[0], t->gemm_tiles[0][0]));
    if (mask & 2u) launch_gemm_group<double, 64, CFG_64>(sg, d, grid(kGridGemm[1], t->gemm_tiles[0][1]));
    if (mask & 4u) launch_gemm_group<double, 128, CFG_128>(sg, d, grid(kGridGemm[2], t->gemm_tiles[0][2]));
    if (mask & 16u) launch_gemm_group<float, 64, CFG_64>(sg, d, grid(kGridGemm[1], t->gemm_tiles[1][1]));
    if (mask & 32u) launch_gemm_group<float, 128, CFG_128>(sg, d, grid(kGridGemm[2], t->gemm_tiles[1][2]));
    if (mask & 64u) {
        hipLaunchKernelGGL(group_gemv_part_kernel, dim3(grid(kGridGemv, t->gemv_units)), dim3(256), 0,
                           sv, d);
        hipLaunchKernelGGL(group_gemv_fold_kernel, dim3(grid(kGridFold, t->fold_units)), dim3(256), 0,
                           sv, d);
    }
    if (mask & 128u)
        hipLaunchKernelGGL(group_trsm_kernel, dim3(grid(kGridTrsm, t->trsm_units)), dim3(1024), 0,
                           st, d);
    if (capture) {
        (void)hipEventRecord(b->ev_join[0], sg);
        (void)hipEventRecord(b->ev_join[1], sv);
        (void)hipEventRecord(b->ev_join[2], st);
        for (int i = 0; i < 3; ++i) (void)hipStreamWaitEvent(s0, b->ev_join[i], 0);
    }
    hipLaunchKernelGGL(publish_kernel, dim3(1), dim3(64), 0, s0, d, b->h_flag);
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
    ok = ok && hipHostMalloc((void **)&b->h_tab, sizeof(BatchTable), hipHostMallocDefault) == hipSuccess;
    ok = ok && hipHostMalloc((void **)&b->h_flag, 64, hipHostMallocDefault) == hipSuccess;
    ok = ok && hipMalloc((void **)&b->d_tab, sizeof(BatchTable)) == hipSuccess;
    ok = ok && hipMalloc((void **)&b->d_ws, kWsDoubles * sizeof(double)) == hipSuccess;
    if (!ok) {
        gpublas_dev_batch_destroy(b);
        return nullptr;
    }
    memset(b->h_tab, 0, sizeof(BatchTable));
    *b->h_flag = 0;
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
    if (b->s) (void)hipStreamDestroy(b->s);
    if (b->h_tab) (void)hipHostFree(b->h_tab);
    if (b->h_flag) (void)hipHostFree(b->h_flag);
    if (b->d_tab) (void)hipFree(b->d_tab);
    if (b->d_ws) (void)hipFree(b->d_ws);
    free(b);
}

void gpublas_dev_batch_clear(gpublas_batch *b) { batch_clear(b); }

int gpublas_dev_batch_count(const gpublas_batch *b)
{
    return b->nd_all + b->ns_all + b->h_tab->ngemv + b->h_tab->ntrsm;
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
    BatchTable *t = b->h_tab;
    if (t->ngemv >= kCap) return 1;
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
        // N: 512-row blocks; column blocks (>= 64 columns) for about 256 units
        const int nrb = ceil_div(m, GV_ROWS);
        int ncb = ceil_div(256, nrb);
        const int maxcb = ceil_div(n, 64);
        if (ncb > maxcb) ncb = maxcb;
        if (ncb < 1) ncb = 1;
        d.span = ceil_div(n, ncb);
        d.split = ceil_div(n, d.span);
        units = nrb * d.split;
        len = m;
    } else {
        // T: 16-column units; row chunks (multiples of 128 rows) for about 256 units
        const int ncb = ceil_div(n, GV_TCOLS);
        int nrc = ceil_div(256, ncb);
        const int maxrc = ceil_div(m, 256);
        if (nrc > maxrc) nrc = maxrc;
        if (nrc < 1) nrc = 1;
        d.span = ceil_div(ceil_div(m, nrc), 128) * 128;
        d.split = ceil_div(m, d.span);
        units = ncb * d.split;
        len = n;
    }
    const size_t need = (size_t)d.split * (size_t)len;
    if (b->ws_used + need > b->ws_cap) return t->ngemv ? 1 : -1;
    d.part = b->d_ws + b->ws_used;
    b->ws_used += (need + 31) / 32 * 32;
    d.unit_begin = t->gemv_units;
    d.fold_begin = t->fold_units;
    t->gemv_units += units;
    t->fold_units += ceil_div(len, GV_FOLD);
    t->gv[t->ngemv++] = d;
    return 0;
}

int gpublas_dev_batch_dtrsm(gpublas_batch *b, char side, char uplo, char transa, char diag, int m,
                            int n, double alpha, const double *d_A, int lda, double *d_B, int ldb)
{
    if (!is_l(side) || m <= 0 || m > TR_MAXM || n <= 0 || alpha == 0.0) return -1;
    BatchTable *t = b->h_tab;
    if (t->ntrsm >= kCap) return 1;
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
    d.unit_begin = t->trsm_units;
    t->trsm_units += ceil_div(n, 16);
    t->tr[t->ntrsm++] = d;
    return 0;
}

int gpublas_dev_batch_launch(gpublas_batch *b)
{
    unsigned mask = 0;
    lay_out(b, &mask);
    b->h_tab->seq = ++b->seq;
    b->launched = 1;
    PROF_BEGIN(b->s, "serve_batch");
    if (b->use_graph) {
        hipGraphExec_t e = graph_for(b, mask);
        if (!e) return (int)hipErrorUnknown;
        const int err = (int)hipGraphLaunch(e, b->s);
        PROF_END(b->s);
        b->graph_launches++;
        return err;
    }
    enqueue(b, mask, false);
    PROF_END(b->s);
    b->eager_launches++;
    return (int)hipGetLastError();
}

int gpublas_dev_batch_done(gpublas_batch *b)
{
    if (!b->launched) return 1;
    return __atomic_load_n(b->h_flag, __ATOMIC_ACQUIRE) >= b->seq;
}

int gpublas_dev_batch_wait(gpublas_batch *b)
{
    if (!b->launched) return 0;
    const int err = (int)hipStreamSynchronize(b->s);
    return err ? err : (gpublas_dev_batch_done(b) ? 0 : (int)hipErrorUnknown);
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
                        unsigned mask = 0;
                        lay_out(b, &mask);
                        b->h_tab->seq = ++b->seq;
                        b->launched = 1;
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
