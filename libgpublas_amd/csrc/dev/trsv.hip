// trsv.hip -- DTRSV on gfx950: solve op(A) x = b in place (replaces
// cublasDtrsv, blas_level2/trsv.cc:81-88 of the reference; netlib semantics),
// and DTPSV, the same solve on the packed triangle.
//
// Orders above 128: one launch (trsv_lookback, below).  The solver
// workgroup of block t applies the last one or two earlier blocks' tiles and
// solves its diagonal block; helper workgroups apply the older tiles as their
// blocks are solved; solutions and partial sums pass between workgroups
// through tagged mailbox words.  0.24 ms at n = 8192 (62 dependent hops of
// about 3.8 us: substitution, block sum, one cross-die round trip).
// Orders up to 128: the same kernel, one solver workgroup, no mailboxes.
// The arithmetic, blocked substitution in 128-wide blocks in solve order
// (forward for lower / no-trans and upper / trans, backward otherwise):
//   * every earlier block's tile of products is subtracted from block t's
//     right-hand side, 16-B loads down the columns of A;
//   * the 128 x 128 diagonal block is solved by substitution in one wave
//     (lane l owns unknowns l and l + 64; the solved value is broadcast with
//     readlane).
// Every element of the triangle is read once.  The order of the sums
// depends only on n, so results repeat bit for bit.
// Inside a diagonal block each row is first multiplied by the reciprocal of
// its diagonal element (r'_i = r_i / d_i, coefficients c(i,k) / d_i, one
// extra rounding each, off the dependent chain), so each of the n dependent
// steps is a readlane and one FMA per row: the chain of 128 steps is fully
// unrolled with immediate lanes and LDS offsets.  Every solved x_k updates
// every later unknown, zero or not (as a GEMV-based blocked solve does), so
// Inf / NaN in the stored triangle propagate to the unknowns solved after
// them, never to those solved before; a zero diagonal element gives
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
// After step K the residual of row K is final: later steps subtract from it
// 0 * x_K' (its coefficients on later unknowns are zero), which leaves it
// unchanged while x_K' is finite.  So the plain chain (CAP = false) leaves
// the solution in r0 / r1 -- unless a solved value is Inf or NaN, where
// 0 * x_K' would spread NaN to rows solved before it.  CAP = true keeps each
// solved value apart as it is found (sol0, sol1), at about twice the
// instructions per step; solve_block runs it only when the plain chain's
// result holds a non-finite value.
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

template <int S, bool FWD, bool CAP>
__device__ __forceinline__ void sub_step(double c0, double c1, double &r0, double &r1,
                                         double &sol0, double &sol1, int lane)
{
    constexpr int K = kth<FWD>(S), KL = K & 63;
    const double xk = readlane_d(K < 64 ? r0 : r1, KL);
    if constexpr (CAP) {
        if constexpr (K < 64) sol0 = lane == KL ? xk : sol0;  // off the dependent chain
        else sol1 = lane == KL ? xk : sol1;
    }
    if constexpr (any_after0<FWD>(K)) r0 = fma(-c0, xk, r0);
    if constexpr (any_after1<FWD>(K)) r1 = fma(-c1, xk, r1);
}

template <bool FWD, bool CAP, int B, int... J>
__device__ __forceinline__ void run_batch(const Coef &c, double &r0, double &r1, double &sol0,
                                          double &sol1, int lane, std::integer_sequence<int, J...>)
{
    (sub_step<B * SB + J, FWD, CAP>(c.c0[J], c.c1[J], r0, r1, sol0, sol1, lane), ...);
}

template <bool FWD, bool CAP, int B>
__device__ __forceinline__ void batches(const double *D, Coef &cur, Coef &nxt, double &r0,
                                        double &r1, double &sol0, double &sol1, int lane)
{
    if constexpr (B < TB / SB) {
        if constexpr (B + 1 < TB / SB)
            load_batch<FWD, B + 1>(D, nxt, lane, std::make_integer_sequence<int, SB>{});
        __builtin_amdgcn_sched_barrier(0);
        run_batch<FWD, CAP, B>(cur, r0, r1, sol0, sol1, lane, std::make_integer_sequence<int, SB>{});
        __builtin_amdgcn_sched_barrier(0);
        batches<FWD, CAP, B + 1>(D, nxt, cur, r0, r1, sol0, sol1, lane);
    }
}

template <bool FWD, bool CAP>
__device__ __forceinline__ void substitute(const double *D, double &r0, double &r1, double &sol0,
                                           double &sol1, int lane)
{
    Coef a, b;
    load_batch<FWD, 0>(D, a, lane, std::make_integer_sequence<int, SB>{});
    batches<FWD, CAP, 0>(D, a, b, r0, r1, sol0, sol1, lane);
}

// the block's solution from its scaled right-hand side r0, r1 (one wave)
template <bool FWD>
__device__ __forceinline__ void solve_block(const double *D, double r0, double r1, double &sol0,
                                            double &sol1, int lane)
{
    double a0 = r0, a1 = r1, u0 = 0.0, u1 = 0.0;
    substitute<FWD, false>(D, a0, a1, u0, u1, lane);
    if (__ballot(!(__builtin_isfinite(a0) && __builtin_isfinite(a1))) == 0) {
        sol0 = a0;
        sol1 = a1;
        return;
    }
    sol0 = sol1 = 0.0;
    substitute<FWD, true>(D, r0, r1, sol0, sol1, lane);
}

// ---- one launch: helpers per block, results through tagged mailboxes --------
// Workgroups come in groups of nh + 1 per solve-order position t: nh helpers,
// then the solver of block t.  Block t's right-hand side needs the products
// of every earlier block's tile (t tiles of 128 x 128 elements); the solver
// applies the last min(2, t) of them itself, helper h the tiles h, h + nh,
// h + 2 nh, ... before those.  So the long rows of a large order are streamed
// by nh + 1 workgroups, tiles are applied as soon as their block is solved,
// and only the solver's last tile sits on the chain of dependences.
// Results travel through mailboxes instead of flags: a value is published as
// two 64-bit words, its low and its high 32 bits each beside the tag 1 in the
// upper half, with relaxed agent-scope atomic stores, and read with relaxed
// agent-scope atomic loads.  A 64-bit atomic word is read whole, so a word
// tagged 1 holds its final bits, and a reader that has seen every tag of a
// block has its values one memory round trip after they were stored, with no
// fence on either side.  The mailboxes are zeroed before the launch.
//   xmail  nb blocks of 128 values: each block's solution (zeros past n)
//   pmail  nb x nh blocks: each helper's sums of products for its row block
// Order of the sums: a helper's lanes accumulate over its tiles in tile
// order, then the per-lane sums are combined in a fixed order (waves in
// order, or three xor-shuffles); the solver subtracts from b the helpers'
// sums in helper order and then the sum of its own tiles.  The order depends
// only on n, so results repeat bit for bit.
// Waits: workgroups of one launch start in index order and every wait is for
// a lower-indexed workgroup (a solver waits for its own helpers, placed
// before it, and for earlier solvers; a helper for earlier solvers), so every
// wait ends; each wait is also bounded in time, and a result that depends on
// an expired wait is NaN, never plausible.
typedef unsigned long long u64;
constexpr int kTrsvHelpers = 3;  // helpers per block (orders of 6+ blocks)
constexpr int kMailWords = 2 * TB;             // words per mailbox block
constexpr u64 kTag = 1ull << 32;

struct TrsvLb {
    int n, nb, nh;     // order, blocks, helpers per block
    const double *A;
    long lda;
    double *x;         // b on entry, the solution on exit
    u64 *xmail;        // nb * kMailWords words
    u64 *pmail;        // nb * nh * kMailWords words
    long long timeout; // steady-clock ticks a wait may last
    int vec;           // A 16-B aligned with even lda
    int packed, plower;  // A is the packed triangle (DTPSV), lower or upper
};

// Column j of A: element (i, j) is colp(g, j)[i].  Packed storage (netlib
// column packing) starts column j at j (j + 1) / 2 (upper) or, shifted by the
// j rows it does not store, at j (2n - j - 1) / 2 (lower).  Inside a diagonal
// block the unstored elements' addresses fall on neighbouring columns, still
// inside the packed array; their values are discarded like full storage's.
__device__ __forceinline__ const double *colp(const TrsvLb &g, long j)
{
    if (!g.packed) return g.A + j * g.lda;
    return g.A + (g.plower ? j * (2 * (long)g.n - j - 1) / 2 : j * (j + 1) / 2);
}

template <bool FWD>
__device__ __forceinline__ void block_of(const TrsvLb &g, int t, int &b0, int &bn)
{
    const int blk = FWD ? t : g.nb - 1 - t;
    b0 = blk * TB;
    bn = min(TB, g.n - b0);
}

__device__ __forceinline__ void mail_put(u64 *m, int i, double v)
{
    const u64 b = __builtin_bit_cast(u64, v);
    __hip_atomic_store(m + 2 * i, kTag | (b & 0xffffffffull), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(m + 2 * i + 1, kTag | (b >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ u64 mail_word(u64 *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool tagged(u64 w) { return (w >> 32) == 1u; }

__device__ __forceinline__ double untag(u64 lo, u64 hi)
{
    return __builtin_bit_cast(double, (hi << 32) | (lo & 0xffffffffull));
}

// The calling wave reads one mailbox block; lane l receives values l and
// l + 64.  light: lane 0 first polls one word alone, so a waiter far from the
// head of the chain loads 8 bytes per poll.  Returns false, with NaN values,
// if the wait outlasted the timeout.
__device__ __forceinline__ bool mail_get(u64 *m, bool light, long long timeout, double &v0,
                                         double &v1)
{
    const int lane = threadIdx.x & 63;
    const long long t0 = wall_clock64();
    int good = 1;
    if (light) {
        if (lane == 0) {
            while (!tagged(mail_word(m + kMailWords - 1))) {
                if (wall_clock64() - t0 > timeout) {
                    good = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        good = __shfl(good, 0, 64);
    }
    u64 w0 = 0, w1 = 0, w2 = 0, w3 = 0;
    while (good) {
        w0 = mail_word(m + 2 * lane);
        w1 = mail_word(m + 2 * lane + 1);
        w2 = mail_word(m + 2 * (lane + 64));
        w3 = mail_word(m + 2 * (lane + 64) + 1);
        const bool ok = tagged(w0) && tagged(w1) && tagged(w2) && tagged(w3);
        if (__ballot(!ok) == 0) break;
        if (wall_clock64() - t0 > timeout) good = 0;
        else __builtin_amdgcn_s_sleep(1);
    }
    if (!good) {
        v0 = v1 = __builtin_nan("");
        return false;
    }
    v0 = untag(w0, w1);
    v1 = untag(w2, w3);
    return true;
}

// One tile's operands in registers: N: lane holds rows c0 + 2 lane, +1 of
// columns p0 + 16 w + u (wave w, u < 16), 16-B loads down each column; T:
// eight lanes (j = lane & 7) share the column segment A[p0 .., c0 + ii] of
// index ii = 16 w + 8 half + (lane >> 3), elements 16 q + 2 j, +1, one
// 128-byte line per eight lanes per load.  Zeros outside the blocks.
struct Tile {
    d2 v[16];
};

template <bool TRANS>
__device__ __forceinline__ void tile_load(const TrsvLb &g, int p0, int pn, int c0, int cn,
                                          Tile &tl)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // whole tiles load unguarded: 16-B loads when A allows them, else
    // (packed columns at any 8-byte offset) two 8-B loads
    const bool whole = pn == TB && cn == TB, full = g.vec && whole;
    if constexpr (!TRANS) {
        const long r = c0 + 2 * lane;
        const int k0 = 16 * wave;
        if (full) {
#pragma unroll
            for (int u = 0; u < 16; ++u)
                tl.v[u] = __builtin_nontemporal_load((const d2 *)(colp(g, p0 + k0 + u) + r));
        } else if (whole) {
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const double *a = colp(g, p0 + k0 + u) + r;
                tl.v[u].x = __builtin_nontemporal_load(a);
                tl.v[u].y = __builtin_nontemporal_load(a + 1);
            }
        } else {
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const bool kk = k0 + u < pn;
                const double *a = colp(g, p0 + k0 + u) + r;
                tl.v[u].x = (kk && 2 * lane < cn) ? a[0] : 0.0;
                tl.v[u].y = (kk && 2 * lane + 1 < cn) ? a[1] : 0.0;
            }
        }
    } else {
        const int c = lane >> 3, j = lane & 7;
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const int ii = 16 * wave + 8 * half + c;
            const double *col = colp(g, c0 + ii) + p0;
#pragma unroll
            for (int q = 0; q < TB / 16; ++q) {
                const int e = 16 * q + 2 * j;
                if (full) {
                    tl.v[8 * half + q] = __builtin_nontemporal_load((const d2 *)(col + e));
                } else if (whole) {
                    tl.v[8 * half + q].x = __builtin_nontemporal_load(col + e);
                    tl.v[8 * half + q].y = __builtin_nontemporal_load(col + e + 1);
                } else {
                    tl.v[8 * half + q].x = (ii < cn && e < pn) ? col[e] : 0.0;
                    tl.v[8 * half + q].y = (ii < cn && e + 1 < pn) ? col[e + 1] : 0.0;
                }
            }
        }
    }
}

// Adds one tile's products to the lane's sums s0, s1 (N: rows 2 lane, +1;
// T: indices of half 0 and 1).  xa / xb: values lane and lane + 64 of the
// tile's column block, as mail_get returns them.
template <bool TRANS>
__device__ __forceinline__ void tile_acc(const Tile &tl, double xa, double xb, double &s0,
                                         double &s1)
{
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if constexpr (!TRANS) {
        const double xs = wave < 4 ? xa : xb;  // x_{16 w + u} is in lane (16 w + u) & 63
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const double xk = readlane_d(xs, (16 * wave + u) & 63);
            s0 = fma(tl.v[u].x, xk, s0);
            s1 = fma(tl.v[u].y, xk, s1);
        }
    } else {
        const int j = lane & 7;
#pragma unroll
        for (int q = 0; q < TB / 16; ++q) {
            const double src = q < 4 ? xa : xb;
            const double xe = __shfl(src, (16 * q + 2 * j) & 63, 64);
            const double xo = __shfl(src, (16 * q + 2 * j + 1) & 63, 64);
            s0 = fma(tl.v[q].x, xe, s0);
            s0 = fma(tl.v[q].y, xo, s0);
            s1 = fma(tl.v[8 + q].x, xe, s1);
            s1 = fma(tl.v[8 + q].y, xo, s1);
        }
    }
}

// The workgroup's lane sums to one sum per index of the row block; wave 0
// receives those of indices lane and lane + 64 in ra, rb.  N: the waves'
// sums in wave order; T: the eight lanes of a column segment by three
// xor-shuffles.  One barrier.
template <bool TRANS>
__device__ __forceinline__ void wg_sums(double s0, double s1, double (*part)[TB], double &ra,
                                        double &rb)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if constexpr (!TRANS) {
        part[wave][2 * lane] = s0;
        part[wave][2 * lane + 1] = s1;
        __syncthreads();
        if (wave == 0) {
            ra = rb = 0.0;
#pragma unroll
            for (int w = 0; w < TTHREADS / 64; ++w) {
                ra += part[w][lane];
                rb += part[w][lane + 64];
            }
        }
    } else {
        const int c = lane >> 3, j = lane & 7;
#pragma unroll
        for (int sh = 1; sh < 8; sh <<= 1) {
            s0 += __shfl_xor(s0, sh, 64);
            s1 += __shfl_xor(s1, sh, 64);
        }
        if (j == 0) {
            part[0][16 * wave + c] = s0;
            part[0][16 * wave + 8 + c] = s1;
        }
        __syncthreads();
        if (wave == 0) {
            ra = part[0][lane];
            rb = part[0][lane + 64];
        }
    }
}

// helper h of block t: the tiles h, h + nh, ... below t - min(2, t), each
// loaded one tile ahead of its wait
template <bool LOWER, bool TRANS>
__device__ void lb_helper(const TrsvLb &g, int t, int h, double (*part)[TB])
{
    constexpr bool FWD = LOWER != TRANS;
    const int last = t - min(2, t);
    if (h >= last) return;
    int c0, cn, p0, pn;
    block_of<FWD>(g, t, c0, cn);
    double s0 = 0.0, s1 = 0.0;
    Tile cur, nxt;
    block_of<FWD>(g, h, p0, pn);
    tile_load<TRANS>(g, p0, pn, c0, cn, cur);
    for (int j = h; j < last; j += g.nh) {
        if (j + g.nh < last) {
            block_of<FWD>(g, j + g.nh, p0, pn);
            tile_load<TRANS>(g, p0, pn, c0, cn, nxt);
        }
        double xa, xb;
        mail_get(g.xmail + (long)j * kMailWords, true, g.timeout, xa, xb);  // NaN if expired
        tile_acc<TRANS>(cur, xa, xb, s0, s1);
        cur = nxt;
    }
    double ra, rb;
    wg_sums<TRANS>(s0, s1, part, ra, rb);
    if (threadIdx.x < 64) {
        u64 *pm = g.pmail + ((long)t * g.nh + h) * kMailWords;
        mail_put(pm, threadIdx.x, ra);
        mail_put(pm, threadIdx.x + 64, rb);
    }
}

// the solver of block t: its diagonal block, the helpers' sums, its own last
// tiles, then the substitution of its diagonal block
template <bool LOWER, bool TRANS, bool UNIT>
__device__ void lb_solver(const TrsvLb &g, int t, double *D, double (*part)[TB], double *rinv)
{
    constexpr bool FWD = LOWER != TRANS;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int c0, cn;
    block_of<FWD>(g, t, c0, cn);
    const bool vec = g.vec != 0;
    const int own = min(2, t), last = t - own;

    // the diagonal block, its diagonal, b and the own tiles (one round trip)
    constexpr int PER = TB * TB / 2 / TTHREADS;
    double va[PER], vb[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int idx = tid + u * TTHREADS;
        const int a = 2 * (idx % (TB / 2)), b = idx / (TB / 2);
        const double *src = colp(g, c0 + b) + (c0 + a);
        if (vec && a + 1 < cn && b < cn) {
            const d2 v = *(const d2 *)src;
            va[u] = v.x;
            vb[u] = v.y;
        } else if (cn == TB) {
            // a whole block: every address is inside A (packed: the unstored
            // half's addresses are the neighbouring columns'), and no select
            // follows the loads (it would wait for each one in turn)
            va[u] = src[0];
            vb[u] = src[1];
        } else {
            va[u] = (a < cn && b < cn) ? src[0] : 0.0;
            vb[u] = (a + 1 < cn && b < cn) ? src[1] : 0.0;
        }
    }
    const double dg = (!UNIT && tid < cn) ? colp(g, c0 + tid)[c0 + tid] : 1.0;
    double acc0 = 0.0, acc1 = 0.0;  // wave 0: b of indices lane, lane + 64, minus the sums
    if (wave == 0) {
        acc0 = lane < cn ? g.x[c0 + lane] : 0.0;
        acc1 = lane + 64 < cn ? g.x[c0 + lane + 64] : 0.0;
    }
    Tile ta, tb;
    int p0, pn;
    if (own >= 1) {
        block_of<FWD>(g, t - own, p0, pn);
        tile_load<TRANS>(g, p0, pn, c0, cn, ta);
    }
    if (own == 2) {
        block_of<FWD>(g, t - 1, p0, pn);
        tile_load<TRANS>(g, p0, pn, c0, cn, tb);
    }
    // Row i of the block's system is divided by its diagonal element d_i
    // (r'_i = r_i / d_i, coefficients c(i,k) / d_i), so the substitution
    // chain has no division.  D[k*LDD + i] holds the scaled coefficient
    // (i, k) of op(A_tt) where unknown i comes after k in solve order and
    // zero elsewhere, so every step updates all rows with one FMA per row.
    // Indices from cn up (a partial block) are padded: zero right-hand side,
    // no coupling.
    if (tid < TB) rinv[tid] = UNIT ? 1.0 : 1.0 / dg;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int idx = tid + u * TTHREADS;
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
    bool good = true;
    if (wave == 0) {
        for (int h = 0; h < g.nh && h < last; ++h) {
            double ra, rb;
            if (!mail_get(g.pmail + ((long)t * g.nh + h) * kMailWords, true, g.timeout, ra, rb))
                good = false;
            acc0 -= ra;
            acc1 -= rb;
        }
    }
    if (own > 0) {
        double s0 = 0.0, s1 = 0.0, xa, xb;
        if (!mail_get(g.xmail + (long)(t - own) * kMailWords, own == 2, g.timeout, xa, xb))
            good = false;
        tile_acc<TRANS>(ta, xa, xb, s0, s1);
        if (own == 2) {
            if (!mail_get(g.xmail + (long)(t - 1) * kMailWords, false, g.timeout, xa, xb))
                good = false;
            tile_acc<TRANS>(tb, xa, xb, s0, s1);
        }
        double ra, rb;
        wg_sums<TRANS>(s0, s1, part, ra, rb);
        if (wave == 0) {
            acc0 -= ra;
            acc1 -= rb;
        }
    } else {
        __syncthreads();  // D
    }
    if (wave != 0) return;
    double r0 = lane < cn ? acc0 * rinv[lane] : 0.0;
    double r1 = lane + 64 < cn ? acc1 * rinv[lane + 64] : 0.0;
    double sol0, sol1;
    solve_block<FWD>(D, r0, r1, sol0, sol1, lane);
    if (!good) sol0 = sol1 = __builtin_nan("");
    if (t + 1 < g.nb) {  // no block reads the last one's mailbox
        u64 *xm = g.xmail + (long)t * kMailWords;
        mail_put(xm, lane, lane < cn ? sol0 : 0.0);
        mail_put(xm, lane + 64, lane + 64 < cn ? sol1 : 0.0);
    }
    if (lane < cn) g.x[c0 + lane] = sol0;
    if (lane + 64 < cn) g.x[c0 + lane + 64] = sol1;
}

template <bool LOWER, bool TRANS, bool UNIT>
__global__ __launch_bounds__(TTHREADS) void trsv_lookback(TrsvLb g)
{
    __shared__ double D[TB * LDD];
    __shared__ double part[TTHREADS / 64][TB];
    __shared__ double rinv[TB];
    const int t = blockIdx.x / (g.nh + 1), role = blockIdx.x % (g.nh + 1);
    if (role < g.nh) lb_helper<LOWER, TRANS>(g, t, role, part);
    else lb_solver<LOWER, TRANS, UNIT>(g, t, D, part, rinv);
}

__global__ void zero_words(u64 *f, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        f[i] = 0ull;
}

// One launch (trsv_lookback) behind a mailbox-clearing kernel; an order up
// to 128 (one block: no mailboxes) is the launch alone.  The mailboxes come
// from the stream's scratch (released behind the stream's work).
template <bool LOWER, bool TRANS, bool UNIT>
int run(hipStream_t s, int n, const double *A, long lda, bool packed, double *x)
{
    TrsvLb g;
    g.n = n;
    g.nb = ceil_div(n, TB);
    g.nh = std::max(0, std::min(kTrsvHelpers, g.nb - 3));
    g.A = A;
    g.lda = lda;
    g.x = x;
    // packed columns start at any 8-byte offset: element loads only
    g.vec = !packed && (((uintptr_t)A & 15) == 0) && (lda % 2 == 0);
    g.packed = packed;
    g.plower = LOWER;
    const long words = g.nb > 1 ? (long)g.nb * (1 + g.nh) * kMailWords : 0;
    g.xmail = g.pmail = nullptr;
    if (words) {
        Scratch scr = dev_scratch((size_t)words * sizeof(u64), s);
        if (!scr.p) return (int)hipErrorOutOfMemory;
        g.xmail = (u64 *)scr.p;
        g.pmail = g.xmail + (long)g.nb * kMailWords;
    }
    int rate_khz = 0;
    if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, 0) != hipSuccess ||
        rate_khz <= 0)
        rate_khz = 100000;
    g.timeout = (long long)rate_khz * 2000;  // 2 s
    const long grid = (long)g.nb * (g.nh + 1);
    if (grid > 0x7fffffffL) return (int)hipErrorInvalidValue;
    if (words) {
        const long zblocks = std::min<long>((words + 255) / 256, 2048);
        hipLaunchKernelGGL(zero_words, dim3((unsigned)zblocks), dim3(256), 0, s, g.xmail, words);
    }
    hipLaunchKernelGGL((trsv_lookback<LOWER, TRANS, UNIT>), dim3((unsigned)grid), dim3(TTHREADS),
                       0, s, g);
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
          bool packed, double *x)
{
    const bool lo = uplo == 'L' || uplo == 'l';
    const bool tr = !(trans == 'N' || trans == 'n');
    const bool un = diag == 'U' || diag == 'u';
#define GB_TRSV(L, T, U) return run<L, T, U>(s, n, A, lda, packed, x)
    if (lo) {
        if (tr) { if (un) GB_TRSV(true, true, true); else GB_TRSV(true, true, false); }
        else { if (un) GB_TRSV(true, false, true); else GB_TRSV(true, false, false); }
    } else {
        if (tr) { if (un) GB_TRSV(false, true, true); else GB_TRSV(false, true, false); }
        else { if (un) GB_TRSV(false, false, true); else GB_TRSV(false, false, false); }
    }
#undef GB_TRSV
}

// x := op(A)^-1 x on full (packed = false) or packed storage, strided x
// through a contiguous copy
int solve_strided(hipStream_t s, char uplo, char trans, char diag, int n, const double *A,
                  long lda, bool packed, double *x, int incx)
{
    if (incx == 1) return solve(s, uplo, trans, diag, n, A, lda, packed, x);
    double *t = (double *)gpublas_dev_alloc((size_t)n * sizeof(double));
    if (!t) return (int)hipErrorOutOfMemory;
    const long x0 = incx < 0 ? -(long)(n - 1) * incx : 0;
    const int blocks = std::min(ceil_div(n, 256), 4096);
    hipLaunchKernelGGL(gather_x, dim3(blocks), dim3(256), 0, s, n, x, (long)incx, x0, t);
    int err = solve(s, uplo, trans, diag, n, A, lda, packed, t);
    if (!err) {
        hipLaunchKernelGGL(scatter_x, dim3(blocks), dim3(256), 0, s, n, t, x, (long)incx, x0);
        err = (int)hipGetLastError();
    }
    dev_release_after(t, s);  // back to the pool once the stream's work on it is done
    return err;
}

}  // namespace

extern "C" int gpublas_dev_dtrsv(void *stream, char uplo, char trans, char diag, int n,
                                 const double *d_A, int lda, double *d_x, int incx)
{
    hipStream_t s = pick_stream(stream);
    if (n <= 0) return 0;
    PROF_BEGIN(s, "dtrsv");
    const int err = solve_strided(s, uplo, trans, diag, n, d_A, lda, false, d_x, incx);
    PROF_END(s);
    return err;
}

// DTPSV (replaces cublasDtpsv behind blas_level2/tpsv.cc): the same one-launch
// solve reading the packed triangle in place
extern "C" int gpublas_dev_dtpsv(void *stream, char uplo, char trans, char diag, int n,
                                 const double *d_AP, double *d_x, int incx)
{
    hipStream_t s = pick_stream(stream);
    if (n <= 0) return 0;
    PROF_BEGIN(s, "dtpsv");
    const int err = solve_strided(s, uplo, trans, diag, n, d_AP, 0, true, d_x, incx);
    PROF_END(s);
    return err;
}
