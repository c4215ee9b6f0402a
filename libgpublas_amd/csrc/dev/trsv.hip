// trsv.hip -- DTRSV on gfx950: solve op(A) x = b in place (replaces
// cublasDtrsv, blas_level2/trsv.cc:81-88 of the reference; netlib semantics).
//
// Orders above 128: one launch (trsv_lookback, below): workgroup t solves
// block t after applying every earlier block as its solution is published
// (per-block flags), 0.47 ms at n = 8192 against 0.8 ms for 64 launches.
// Orders up to 128, and the description of the arithmetic both paths share:
// blocked right-looking substitution, one launch per 128-wide block in solve
// order (forward for lower / no-trans and upper / trans, backward
// otherwise).  Launch t applies the block solved by launch t-1 to every
// unknown still open and solves block t:
//   * workgroup 0 applies it to block t's own 128 unknowns, then solves the
//     128 x 128 diagonal block by substitution in one wave (lane l owns
//     unknowns l and l + 64; the solved value is broadcast with readlane);
//   * the other workgroups apply it to the open unknowns after block t,
//     streaming the panel once with 16-B loads.
// Every element of the triangle is read once, by one launch.  The update of
// an unknown by block j happens in launch j + 1, so the order of the sums
// depends only on n (results repeat bit for bit); within a launch the 128
// products of a block are summed in a fixed order.
// Inside a diagonal block each row is first multiplied by the reciprocal of
// its diagonal element (r'_i = r_i / d_i, coefficients c(i,k) / d_i, one
// extra rounding each, off the dependent chain), so each of the n dependent
// steps is a readlane and one FMA per row: the chain of 128 steps is fully
// unrolled with immediate lanes and LDS offsets.  Every solved x_k updates
// every later unknown, zero or not (as a GEMV-based blocked solve does), so
// Inf / NaN in the stored triangle propagate; a zero diagonal element gives
// non-finite results.
//
#include "devcommon.hpp"
#include "gpublas_dev.h"
#include <algorithm>
#include <utility>

namespace {

constexpr int TB = 128;           // block width
constexpr int TTHREADS = 512;     // 8 waves per workgroup
constexpr int LDD = TB + 1;       // LDS leading dimension of the diagonal block (doubles)

typedef double d2 __attribute__((ext_vector_type(2)));

// Sum over the previous block's 128 indices for 128 consecutive open
// indices i0.. of the update
//   N: r_i -= sum_k A[i, k] x_k     (A rows i, columns k of the previous block)
//   T: r_i -= sum_k A[k, i] x_k     (A rows k of the previous block, column i)
// by one 512-thread workgroup; returns the 128 sums in red[] (LDS) in a
// fixed order.  N: wave w sums columns 16w..16w+15 for all 128 rows (lane:
// 2 rows, one 16-B load per column), partials added over the waves in order.
// T: wave w owns indices 16w..16w+15, eight lanes per column segment (128
// doubles), reduced across those lanes.
template <bool TRANS>
__device__ void block_products(int n, const double *A, long lda, const double *xp, int p0, int pn,
                               long i0, int in, bool vec, double *red, double (*part)[TB])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if constexpr (!TRANS) {
        const long r = i0 + 2 * lane;
        double s0 = 0.0, s1 = 0.0;
        const int k0 = 16 * wave, k1 = min(pn, k0 + 16);
        if (vec && 2 * lane + 1 < in) {
            d2 v[16];
#pragma unroll
            for (int u = 0; u < 16; ++u)
                if (k0 + u < k1) v[u] = *(const d2 *)(A + r + (long)(p0 + k0 + u) * lda);
#pragma unroll
            for (int u = 0; u < 16; ++u)
                if (k0 + u < k1) {
                    const double xk = xp[k0 + u];
                    s0 = fma(v[u].x, xk, s0);
                    s1 = fma(v[u].y, xk, s1);
                }
        } else {
            for (int k = k0; k < k1; ++k) {
                const double xk = xp[k];
                if (2 * lane < in) s0 = fma(A[r + (long)(p0 + k) * lda], xk, s0);
                if (2 * lane + 1 < in) s1 = fma(A[r + 1 + (long)(p0 + k) * lda], xk, s1);
            }
        }
        part[wave][2 * lane] = s0;
        part[wave][2 * lane + 1] = s1;
        __syncthreads();
        if (threadIdx.x < TB) {
            double t = 0.0;
#pragma unroll
            for (int w = 0; w < TTHREADS / 64; ++w) t += part[w][threadIdx.x];
            red[threadIdx.x] = t;
        }
        __syncthreads();
    } else {
        // column segment of index i: A[p0 .. p0+pn-1, i].  Eight lanes share a
        // column (lane = 8c + j): pass q reads elements 16q + 2j, 16q + 2j + 1,
        // so the eight lanes of a column read one 128-byte line per
        // instruction; their partial sums are added by three xor-shuffles.
        const int c = lane >> 3, j = lane & 7;
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const int ii = 16 * wave + 8 * half + c;
            const double *col = A + p0 + (i0 + ii) * lda;
            double t = 0.0;
            if (ii < in) {
                if (vec && pn == TB) {
                    d2 v[TB / 16];
#pragma unroll
                    for (int q = 0; q < TB / 16; ++q) v[q] = *(const d2 *)(col + 16 * q + 2 * j);
#pragma unroll
                    for (int q = 0; q < TB / 16; ++q) {
                        t = fma(v[q].x, xp[16 * q + 2 * j], t);
                        t = fma(v[q].y, xp[16 * q + 2 * j + 1], t);
                    }
                } else {
                    for (int e = 2 * j; e < pn; e += 16) {
                        t = fma(col[e], xp[e], t);
                        if (e + 1 < pn) t = fma(col[e + 1], xp[e + 1], t);
                    }
                }
            }
            t += __shfl_xor(t, 1, 64);
            t += __shfl_xor(t, 2, 64);
            t += __shfl_xor(t, 4, 64);
            if (j == 0 && ii < TB) red[ii] = t;
        }
        __syncthreads();
    }
}

__device__ __forceinline__ double readlane_d(double v, int l)
{
    const long b = __builtin_bit_cast(long, v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __builtin_bit_cast(double, ((long)hi << 32) | (unsigned)lo);
}

// The substitution: step S solves unknown K = kth(S) of the block (S, K
// compile-time constants, so lanes and LDS offsets are immediates); r0 / r1
// hold the scaled residuals of unknowns lane and lane + 64.  A step is a
// readlane of x_K and one FMA per row.  The coefficients come from LDS in
// batches of SB steps, each batch's loads issued while the previous batch
// runs (scheduling barriers keep the loads ahead of the chain).
constexpr int SB = 16;

template <bool FWD> __host__ __device__ constexpr int kth(int s) { return FWD ? s : TB - 1 - s; }
// unknowns after K in solve order among lanes' rows 0..63 and 64..127
template <bool FWD> __host__ __device__ constexpr bool any_after0(int k) { return FWD ? k < 63 : k > 0; }
template <bool FWD> __host__ __device__ constexpr bool any_after1(int k) { return FWD ? k < 127 : k > 64; }

struct Coef {
    double c0[SB], c1[SB];
};

template <bool FWD, int B, int... J>
__device__ __forceinline__ void load_batch(const double *D, Coef &c, int lane,
                                           std::integer_sequence<int, J...>)
{
    ((c.c0[J] = any_after0<FWD>(kth<FWD>(B * SB + J)) ? D[kth<FWD>(B * SB + J) * LDD + lane] : 0.0),
     ...);
    ((c.c1[J] = any_after1<FWD>(kth<FWD>(B * SB + J)) ? D[kth<FWD>(B * SB + J) * LDD + lane + 64]
                                                       : 0.0),
     ...);
}

template <int S, bool FWD>
__device__ __forceinline__ void sub_step(double c0, double c1, double &r0, double &r1,
                                         double &sol0, double &sol1, int lane)
{
    constexpr int K = kth<FWD>(S), KL = K & 63;
    const double xk = readlane_d(K < 64 ? r0 : r1, KL);
    if constexpr (K < 64) sol0 = lane == KL ? xk : sol0;  // off the dependent chain
    else sol1 = lane == KL ? xk : sol1;
    if constexpr (any_after0<FWD>(K)) r0 = fma(-c0, xk, r0);
    if constexpr (any_after1<FWD>(K)) r1 = fma(-c1, xk, r1);
}

template <bool FWD, int B, int... J>
__device__ __forceinline__ void run_batch(const Coef &c, double &r0, double &r1, double &sol0,
                                          double &sol1, int lane, std::integer_sequence<int, J...>)
{
    (sub_step<B * SB + J, FWD>(c.c0[J], c.c1[J], r0, r1, sol0, sol1, lane), ...);
}

template <bool FWD, int B>
__device__ __forceinline__ void batches(const double *D, Coef &cur, Coef &nxt, double &r0,
                                        double &r1, double &sol0, double &sol1, int lane)
{
    if constexpr (B < TB / SB) {
        if constexpr (B + 1 < TB / SB)
            load_batch<FWD, B + 1>(D, nxt, lane, std::make_integer_sequence<int, SB>{});
        __builtin_amdgcn_sched_barrier(0);
        run_batch<FWD, B>(cur, r0, r1, sol0, sol1, lane, std::make_integer_sequence<int, SB>{});
        __builtin_amdgcn_sched_barrier(0);
        batches<FWD, B + 1>(D, nxt, cur, r0, r1, sol0, sol1, lane);
    }
}

template <bool FWD>
__device__ __forceinline__ void substitute(const double *D, double &r0, double &r1, double &sol0,
                                           double &sol1, int lane)
{
    Coef a, b;
    load_batch<FWD, 0>(D, a, lane, std::make_integer_sequence<int, SB>{});
    batches<FWD, 0>(D, a, b, r0, r1, sol0, sol1, lane);
}

struct TrsvStep {
    int n, nb;         // order, number of blocks
    const double *A;
    long lda;
    double *x;         // contiguous right-hand side / solution
    int t;             // solve-order position of the block solved here
    int vec;           // A 16-B aligned with even lda
};

// block at solve-order position t: [b0, b0 + bn)
template <bool FWD>
__device__ __forceinline__ void block_at(const TrsvStep &g, int t, int &b0, int &bn)
{
    const int blk = FWD ? t : g.nb - 1 - t;
    b0 = blk * TB;
    bn = min(TB, g.n - b0);
}

template <bool LOWER, bool TRANS, bool UNIT>
__global__ __launch_bounds__(TTHREADS) void trsv_step(TrsvStep g)
{
    constexpr bool FWD = LOWER != TRANS;
    __shared__ double D[TB * LDD];        // diagonal block, D[k * LDD + i] = coefficient (i, k)
    __shared__ double part[TTHREADS / 64][TB];
    __shared__ double red[TB];
    __shared__ double xp[TB];             // the previous block's solution
    __shared__ double rinv[TB];           // 1 / d_i (1 for a unit diagonal)
    int c0, cn;
    block_at<FWD>(g, g.t, c0, cn);
    int p0 = 0, pn = 0;
    if (g.t > 0) block_at<FWD>(g, g.t - 1, p0, pn);
    if (g.t > 0 && threadIdx.x < pn) xp[threadIdx.x] = g.x[p0 + threadIdx.x];
    const bool vec = g.vec != 0;

    if (blockIdx.x > 0) {
        // ---- open unknowns after block t: 128 per workgroup ----
        if (g.t == 0) return;
        long lo, hi;  // open range after the current block, in solve order
        if (FWD) { lo = (long)c0 + cn; hi = g.n; }
        else { lo = 0; hi = c0; }
        const long i0 = lo + (long)(blockIdx.x - 1) * TB;
        if (i0 >= hi) return;
        const int in = (int)min((long)TB, hi - i0);
        __syncthreads();  // xp
        block_products<TRANS>(g.n, g.A, g.lda, xp, p0, pn, i0, in, vec, red, part);
        if (threadIdx.x < in) g.x[i0 + threadIdx.x] -= red[threadIdx.x];
        return;
    }

    // ---- workgroup 0: the diagonal block of position t ----
    // Row i of the block's system is divided by its diagonal element d_i
    // first (r'_i = r_i / d_i, coefficients c(i,k) / d_i), so the
    // substitution chain itself has no division.  D[k*LDD + i] holds the
    // scaled coefficient (i, k) of op(A_cc) where unknown i comes after k in
    // solve order and zero elsewhere, so every step updates all rows with
    // one FMA per row; the solved values are kept apart (sol0, sol1) and
    // the rows already solved are never read again.  Indices from cn up
    // (a partial last block) are padded: zero right-hand side, no coupling.
    // The loads of the diagonal block, its diagonal, the block's right-hand
    // side and the previous block's panel tile are all issued before the
    // first wait, so the workgroup pays one memory round trip.
    constexpr int PER = TB * TB / 2 / TTHREADS;  // row pairs per thread
    double va[PER], vb[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int idx = threadIdx.x + u * TTHREADS;
        const int a = 2 * (idx % (TB / 2)), b = idx / (TB / 2);
        const double *src = g.A + (c0 + a) + (long)(c0 + b) * g.lda;
        if (vec && a + 1 < cn && b < cn) {
            const d2 v = *(const d2 *)src;
            va[u] = v.x;
            vb[u] = v.y;
        } else {
            va[u] = (a < cn && b < cn) ? src[0] : 0.0;
            vb[u] = (a + 1 < cn && b < cn) ? src[1] : 0.0;
        }
    }
    const int tid = threadIdx.x;
    const double dg = (!UNIT && tid < cn) ? g.A[(c0 + tid) + (long)(c0 + tid) * g.lda] : 1.0;
    const double xr = tid < cn ? g.x[c0 + tid] : 0.0;
    if (g.t > 0) block_products<TRANS>(g.n, g.A, g.lda, xp, p0, pn, c0, cn, vec, red, part);
    if (tid < TB) rinv[tid] = UNIT ? 1.0 : 1.0 / dg;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int idx = threadIdx.x + u * TTHREADS;
        const int a0 = 2 * (idx % (TB / 2)), b = idx / (TB / 2);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int a = a0 + h;  // element A[c0 + a, c0 + b]
            const int i = TRANS ? b : a, k = TRANS ? a : b;  // its coefficient (i, k)
            const bool after = FWD ? i > k : i < k;
            double v = h ? vb[u] : va[u];
            v = (after && i < cn && k < cn) ? (UNIT ? v : v * rinv[i]) : 0.0;
            D[k * LDD + i] = v;
        }
    }
    // scaled right-hand side of the block (zero beyond cn), through LDS
    if (tid < TB) xp[tid] = tid < cn ? (xr - (g.t > 0 ? red[tid] : 0.0)) * rinv[tid] : 0.0;
    __syncthreads();
    if (tid >= 64) return;
    const int lane = tid;
    double r0 = xp[lane], r1 = xp[lane + 64];
    double sol0 = 0.0, sol1 = 0.0;
    substitute<FWD>(D, r0, r1, sol0, sol1, lane);
    if (lane < cn) g.x[c0 + lane] = sol0;
    if (lane + 64 < cn) g.x[c0 + lane + 64] = sol1;
}

// ---- one launch: every block a workgroup, ordered by flags ----------------
// Workgroup t (solve-order position t) applies the blocks solved before it as
// their solutions appear -- a flag per block in device memory, set with a
// release after the block's x is stored and read with an acquire -- and then
// solves its own diagonal block as trsv_step's workgroup 0 does.  The older
// blocks are applied while the workgroup would wait anyway; the panel tile of
// the last one (the only one on the dependent chain) is loaded into
// registers before its flag is awaited.  Sums are formed and subtracted in
// the same order as by the launch-per-block path (x_i = ((b_i - r_0) - r_1)
// - ..., each r_j summed in block_products' fixed order), so both paths give
// the same bits.  Workgroups start in index order, and a workgroup waits only
// for lower-indexed ones, so every wait ends; each wait is also bounded in
// time (a solution block that would have come from an expired wait is set to
// NaN, never left plausible).
struct TrsvLb {
    int n, nb;
    const double *A;
    long lda;
    double *x;
    unsigned *flags;   // nb words, zero at launch
    long long timeout; // steady-clock ticks a wait may last
    int vec;
};

// The panel tile of one block (all 128 x 128 products' operands of one
// workgroup) in registers, the loads of block_products without the sums.
struct PanelRegs {
    d2 v[16];
};

template <bool TRANS>
__device__ __forceinline__ void panel_load(const TrsvLb &g, int p0, long i0, PanelRegs &pr)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if constexpr (!TRANS) {
        const long r = i0 + 2 * lane;
        const int k0 = 16 * wave;
#pragma unroll
        for (int u = 0; u < 16; ++u)
            pr.v[u] = __builtin_nontemporal_load((const d2 *)(g.A + r + (long)(p0 + k0 + u) * g.lda));
    } else {
        const int c = lane >> 3, j = lane & 7;
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const int ii = 16 * wave + 8 * half + c;
            const double *col = g.A + p0 + (i0 + ii) * g.lda;
#pragma unroll
            for (int q = 0; q < TB / 16; ++q)
                pr.v[8 * half + q] = __builtin_nontemporal_load((const d2 *)(col + 16 * q + 2 * j));
        }
    }
}

// the sums of block_products (full vec tile) from registers; red[] gets the
// 128 results
template <bool TRANS>
__device__ __forceinline__ void panel_apply(const PanelRegs &pr, const double *xp, double *red,
                                            double (*part)[TB])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if constexpr (!TRANS) {
        double s0 = 0.0, s1 = 0.0;
        const int k0 = 16 * wave;
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const double xk = xp[k0 + u];
            s0 = fma(pr.v[u].x, xk, s0);
            s1 = fma(pr.v[u].y, xk, s1);
        }
        part[wave][2 * lane] = s0;
        part[wave][2 * lane + 1] = s1;
        __syncthreads();
        if (threadIdx.x < TB) {
            double t = 0.0;
#pragma unroll
            for (int w = 0; w < TTHREADS / 64; ++w) t += part[w][threadIdx.x];
            red[threadIdx.x] = t;
        }
        __syncthreads();
    } else {
        const int c = lane >> 3, j = lane & 7;
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const int ii = 16 * wave + 8 * half + c;
            double t = 0.0;
#pragma unroll
            for (int q = 0; q < TB / 16; ++q) {
                t = fma(pr.v[8 * half + q].x, xp[16 * q + 2 * j], t);
                t = fma(pr.v[8 * half + q].y, xp[16 * q + 2 * j + 1], t);
            }
            t += __shfl_xor(t, 1, 64);
            t += __shfl_xor(t, 2, 64);
            t += __shfl_xor(t, 4, 64);
            if (j == 0) red[ii] = t;
        }
        __syncthreads();
    }
}

// thread 0 waits for block j's flag; false if the wait expired.  The polls
// are relaxed loads (an acquire per poll would invalidate the cache for every
// other workgroup streaming its panel); one acquire fence follows.
__device__ __forceinline__ bool wait_block(const TrsvLb &g, int j, int *ok)
{
    if (threadIdx.x == 0) {
        const long long t0 = wall_clock64();
        int good = 1;
        while (__hip_atomic_load(g.flags + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
            if (wall_clock64() - t0 > g.timeout) {
                good = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        *ok = good;
    }
    __syncthreads();
    return *ok != 0;
}

template <bool LOWER, bool TRANS, bool UNIT>
__global__ __launch_bounds__(TTHREADS) void trsv_lookback(TrsvLb g)
{
    constexpr bool FWD = LOWER != TRANS;
    __shared__ double D[TB * LDD];
    __shared__ double part[TTHREADS / 64][TB];
    __shared__ double red[TB];
    __shared__ double xp[TB];
    __shared__ double rinv[TB];
    __shared__ int ok;
    const int t = blockIdx.x;
    const int blk = FWD ? t : g.nb - 1 - t;
    const int c0 = blk * TB, cn = min(TB, g.n - c0);
    const int tid = threadIdx.x;
    const bool vec = g.vec != 0;

    // the diagonal block, its diagonal and the right-hand side (one round trip)
    constexpr int PER = TB * TB / 2 / TTHREADS;
    double va[PER], vb[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int idx = tid + u * TTHREADS;
        const int a = 2 * (idx % (TB / 2)), b = idx / (TB / 2);
        const double *src = g.A + (c0 + a) + (long)(c0 + b) * g.lda;
        if (vec && a + 1 < cn && b < cn) {
            const d2 v = *(const d2 *)src;
            va[u] = v.x;
            vb[u] = v.y;
        } else {
            va[u] = (a < cn && b < cn) ? src[0] : 0.0;
            vb[u] = (a + 1 < cn && b < cn) ? src[1] : 0.0;
        }
    }
    const double dg = (!UNIT && tid < cn) ? g.A[(c0 + tid) + (long)(c0 + tid) * g.lda] : 1.0;
    double acc = tid < cn ? g.x[c0 + tid] : 0.0;  // b_i, then minus each block's sums
    if (tid < TB) rinv[tid] = UNIT ? 1.0 : 1.0 / dg;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int idx = tid + u * TTHREADS;
        const int a0 = 2 * (idx % (TB / 2)), b = idx / (TB / 2);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int a = a0 + h;
            const int i = TRANS ? b : a, k = TRANS ? a : b;
            const bool after = FWD ? i > k : i < k;
            double v = h ? vb[u] : va[u];
            v = (after && i < cn && k < cn) ? (UNIT ? v : v * rinv[i]) : 0.0;
            D[k * LDD + i] = v;
        }
    }
    bool good = true;
    const bool full = vec && cn == TB;
    // blocks solved before this one, in solve order
    for (int j = 0; j < t; ++j) {
        const int pb = FWD ? j : g.nb - 1 - j;
        const int p0 = pb * TB, pn = min(TB, g.n - p0);
        const bool reg = full && pn == TB;
        PanelRegs pr;
        if (reg) panel_load<TRANS>(g, p0, c0, pr);  // before the wait
        if (!wait_block(g, j, &ok)) good = false;
        if (tid < pn) xp[tid] = g.x[p0 + tid];
        __syncthreads();
        if (reg) panel_apply<TRANS>(pr, xp, red, part);
        else block_products<TRANS>(g.n, g.A, g.lda, xp, p0, pn, c0, cn, vec, red, part);
        if (tid < cn) acc -= red[tid];
    }
    if (tid < TB) xp[tid] = tid < cn ? acc * rinv[tid] : 0.0;
    __syncthreads();
    if (tid < 64) {
        const int lane = tid;
        double r0 = xp[lane], r1 = xp[lane + 64];
        double sol0 = 0.0, sol1 = 0.0;
        substitute<FWD>(D, r0, r1, sol0, sol1, lane);
        if (!good) sol0 = sol1 = __builtin_nan("");
        if (lane < cn) g.x[c0 + lane] = sol0;
        if (lane + 64 < cn) g.x[c0 + lane + 64] = sol1;
        __threadfence();  // x of this block before its flag, for every workgroup
        if (lane == 0) __hip_atomic_store(g.flags + t, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ void zero_flags(unsigned *f, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) f[i] = 0u;
}

// One launch (trsv_lookback) behind a flag-clearing kernel.  Its flags come
// from the stream's scratch (released behind the stream's work).
template <bool LOWER, bool TRANS, bool UNIT>
int run_lookback(hipStream_t s, int n, const double *A, long lda, double *x)
{
    TrsvLb g;
    g.n = n;
    g.nb = ceil_div(n, TB);
    g.A = A;
    g.lda = lda;
    g.x = x;
    g.vec = (((uintptr_t)A & 15) == 0) && (lda % 2 == 0);
    Scratch scr = dev_scratch((size_t)g.nb * sizeof(unsigned), s);
    if (!scr.p) return (int)hipErrorOutOfMemory;
    g.flags = (unsigned *)scr.p;
    int rate_khz = 0;
    if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, 0) != hipSuccess ||
        rate_khz <= 0)
        rate_khz = 100000;
    g.timeout = (long long)rate_khz * 2000;  // 2 s
    hipLaunchKernelGGL(zero_flags, dim3((unsigned)std::min(ceil_div(g.nb, 256), 1024)), dim3(256),
                       0, s, g.flags, g.nb);
    hipLaunchKernelGGL((trsv_lookback<LOWER, TRANS, UNIT>), dim3((unsigned)g.nb), dim3(TTHREADS),
                       0, s, g);
    return (int)hipGetLastError();
}

template <bool LOWER, bool TRANS, bool UNIT>
int run(hipStream_t s, int n, const double *A, long lda, double *x)
{
    if (n > TB) return run_lookback<LOWER, TRANS, UNIT>(s, n, A, lda, x);
    TrsvStep g;
    g.n = n;
    g.nb = ceil_div(n, TB);
    g.A = A;
    g.lda = lda;
    g.x = x;
    g.vec = (((uintptr_t)A & 15) == 0) && (lda % 2 == 0);
    constexpr bool FWD = LOWER != TRANS;
    for (int t = 0; t < g.nb; ++t) {
        g.t = t;
        // workgroup 0 plus one per 128 open unknowns after block t (forward:
        // the blocks after it, the last one possibly partial; backward: the
        // full blocks before it)
        const int open = FWD ? n - std::min(n, (t + 1) * TB) : (g.nb - 1 - t) * TB;
        const int blocks = 1 + (t > 0 && open > 0 ? ceil_div(open, TB) : 0);
        hipLaunchKernelGGL((trsv_step<LOWER, TRANS, UNIT>), dim3((unsigned)blocks),
                           dim3(TTHREADS), 0, s, g);
    }
    return (int)hipGetLastError();
}

__global__ void gather_x(int n, const double *x, long incx, long x0, double *t)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (long)gridDim.x * blockDim.x) t[i] = x[x0 + i * incx];
}

__global__ void scatter_x(int n, const double *t, double *x, long incx, long x0)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (long)gridDim.x * blockDim.x) x[x0 + i * incx] = t[i];
}

int solve(hipStream_t s, char uplo, char trans, char diag, int n, const double *A, long lda,
          double *x)
{
    const bool lo = uplo == 'L' || uplo == 'l';
    const bool tr = !(trans == 'N' || trans == 'n');
    const bool un = diag == 'U' || diag == 'u';
#define GB_TRSV(L, T, U) return run<L, T, U>(s, n, A, lda, x)
    if (lo) {
        if (tr) { if (un) GB_TRSV(true, true, true); else GB_TRSV(true, true, false); }
        else { if (un) GB_TRSV(true, false, true); else GB_TRSV(true, false, false); }
    } else {
        if (tr) { if (un) GB_TRSV(false, true, true); else GB_TRSV(false, true, false); }
        else { if (un) GB_TRSV(false, false, true); else GB_TRSV(false, false, false); }
    }
#undef GB_TRSV
}

}  // namespace

extern "C" int gpublas_dev_dtrsv(void *stream, char uplo, char trans, char diag, int n,
                                 const double *d_A, int lda, double *d_x, int incx)
{
    hipStream_t s = pick_stream(stream);
    if (n <= 0) return 0;
    PROF_BEGIN(s, "dtrsv");
    int err;
    if (incx == 1) {
        err = solve(s, uplo, trans, diag, n, d_A, lda, d_x);
    } else {
        double *t = (double *)gpublas_dev_alloc((size_t)n * sizeof(double));
        if (!t) return (int)hipErrorOutOfMemory;
        const long x0 = incx < 0 ? -(long)(n - 1) * incx : 0;
        const int blocks = std::min(ceil_div(n, 256), 4096);
        hipLaunchKernelGGL(gather_x, dim3(blocks), dim3(256), 0, s, n, d_x, (long)incx, x0, t);
        err = solve(s, uplo, trans, diag, n, d_A, lda, t);
        if (!err) {
            hipLaunchKernelGGL(scatter_x, dim3(blocks), dim3(256), 0, s, n, t, d_x, (long)incx, x0);
            err = (int)hipGetLastError();
        }
        dev_release_after(t, s);  // back to the pool once the stream's work on it is done
    }
    PROF_END(s);
    return err;
}
