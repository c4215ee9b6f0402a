// ozaki.hip -- opt-in split-precision DGEMM / DSYRK on BF16 MFMA.
//
// The error-free splitting of Ozaki, Ogita, Oishi & Rump ("Error-free
// transformations of matrix multiplication by using fast routines of matrix
// multiplication and its applications", Numer. Algorithms 59, 2012), in the
// integer-slice form:
//  1. every row i of op(A) is scaled by 2^-ea_i and every column j of op(B)
//     by 2^-eb_j so its largest entry has magnitude below 1 (exact);
//  2. each scaled entry x is cut into s signed 8-bit integers:
//       t = 256 x;  q_p = trunc(t);  x = t - q_p      (p = 0..s-1, all exact)
//     so a = 2^(ea-8) sum_p q_p 2^-8p + r with |r| < 2^(ea-8s).  Slice p is
//     stored as q_p 2^-8p, exact in BF16 (|q_p| <= 255: 8 significant bits);
//  3. the slice products with p + q <= s - 1 run on v_mfma_f32_32x32x16_bf16
//     and accumulate in FP32 by diagonal d = p + q.  Diagonals 0 to 3 are
//     exact over k chunks of 256, 128, 64 and 64 (every partial sum is a
//     multiple of 2^-8d below 2^(24-8d)); diagonal 4 and the diagonals
//     from 5 on, whose weight is below 2^-32 of the leading term, have one
//     rounding accumulator each;
//  4. each accumulator is added into FP64 at the end of its chunk; the
//     epilogue applies 2^(ea_i + eb_j - 16), alpha and beta in FP64.
// The approximations are the dropped pairs (p + q >= s), the slice
// remainders, the FP32 rounding of the tail accumulators and the FP64
// additions.
// FP64 is used exactly where the scheme needs it: the scaling/splitting pass
// and the final summation.
#include "devcommon.hpp"
#include "gpublas_dev.h"
#include <math.h>
#include <algorithm>
#include <type_traits>

namespace {

typedef __bf16 bf16;
typedef short v8s __attribute__((ext_vector_type(8)));  // 16 B of bf16 for loads
constexpr int OZ_MAXS = 8;
// Product tiles are 64 columns wide with one 32 x 32 MFMA tile per wave,
// which keeps a wave's accumulators under 256 registers so two waves share
// each SIMD.  Up to four slices: 64-row tiles, 4 waves, two workgroups per CU
// (independent barriers); five or more: 128-row tiles, 8 waves, one
// workgroup per CU, which moves a quarter fewer bytes from L2 per flop.
constexpr int oz_tile_rows(int s) { return s <= 4 ? 64 : 128; }

struct OzArgs {
    int M, N, Mp, Np, KT;   // Mp, Np: rows padded to 128; KT: 16-deep k slabs
    const bf16 *SA, *SB;    // [s][KT][Mp][16], [s][KT][Np][16]
    const int *ea, *eb;     // row exponents of op(A), column exponents of op(B)
    double alpha, beta;
    double *C;
    long ldc;
    int tiles_m, tiles_n;
    int tri;                // 0 rectangle, 1 lower, 2 upper (SYRK)
};

// ---- exponents -----------------------------------------------------------
// e such that max|x| < 2^e (0 for an all-zero line); a non-finite entry sets
// *bad so the caller can fall back to plain FP64.
__device__ __forceinline__ int exp_of(double mx)
{
    if (mx == 0.0) return 0;
    int e;
    frexp(mx, &e);  // mx = f * 2^e, f in [0.5, 1)
    return e;
}

// line r of an operand stored so that consecutive r are contiguous
// (element (r, k) at X[r + k*ld]): thread per line, coalesced over r, with K
// cut into gridDim.y chunks so the pass fills the chip; chunk c writes the
// partial maximum of line r to part[c*R + r].
__global__ void exp_rows_strided(int R, int K, const double *X, long ld, double *part, int *bad)
{
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const int chunk = (K + gridDim.y - 1) / gridDim.y;
    const int k0 = blockIdx.y * chunk, k1 = min(K, k0 + chunk);
    double mx = 0;
    for (int k = k0; k < k1; ++k) {
        const double v = fabs(X[r + (long)k * ld]);
        if (!(v <= 1.7976931348623157e308)) *bad = 1;
        mx = fmax(mx, v);
    }
    part[(long)blockIdx.y * R + r] = mx;
}

__global__ void exp_rows_finish(int R, int chunks, const double *part, int *e)
{
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    double mx = 0;
    for (int c = 0; c < chunks; ++c) mx = fmax(mx, part[(long)c * R + r]);
    e[r] = exp_of(mx);
}

// line r stored contiguously (element (r, k) at X[k + r*ld]): one wave per line.
__global__ void exp_rows_contig(int R, int K, const double *X, long ld, int *e, int *bad)
{
    const int lane = threadIdx.x & 63;
    const long r = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= R) return;
    double mx = 0;
    for (int k = lane; k < K; k += 64) {
        const double v = fabs(X[k + r * ld]);
        if (!(v <= 1.7976931348623157e308)) *bad = 1;
        mx = fmax(mx, v);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if (lane == 0) e[r] = exp_of(mx);
}

// ---- splitting -----------------------------------------------------------
// Slices are stored pre-scaled, slice p (0-based) holding q_p * 2^-8p, so a
// product of slices p and q carries its true relative weight 2^-8(p+q) and
// the products of one "diagonal" d = p + q can share an accumulator.  Layout
// S[p][k/16][row][16] (bf16, rows padded to a multiple of 128, k to 16): a
// 16-deep k slab of 32 consecutive rows is one contiguous kilobyte, the unit
// the product kernel moves straight from HBM into LDS.
// 32 (r) x 64 (k) tile through LDS so both the FP64 read (either layout) and
// the slice writes are coalesced.
__global__ __launch_bounds__(256) void split_kernel(int R, int Rp, int K, int KT, int s,
                                                    const double *X, long ld, int rcontig,
                                                    const int *e, int f16, bf16 *S)
{
    __shared__ double t[32][65];
    const int r0 = blockIdx.x * 32, k0 = blockIdx.y * 64;
    for (int idx = threadIdx.x; idx < 32 * 64; idx += 256) {
        int rr, kk;
        if (rcontig) { rr = idx % 32; kk = idx / 32; }  // r fastest: X[r + k*ld]
        else { kk = idx % 64; rr = idx / 64; }          // k fastest: X[k + r*ld]
        const long r = r0 + rr, k = k0 + kk;
        double v = 0;
        if (r < R && k < K) v = rcontig ? X[r + k * ld] : X[k + r * ld];
        t[rr][kk] = v;
    }
    __syncthreads();
    // thread -> (k slab kb, row, k pair): 8 threads write one row's 32 bytes
    for (int idx = threadIdx.x; idx < 4 * 32 * 8; idx += 256) {
        const int kb = idx / 256, rr = (idx % 256) / 8, kp = idx % 8;
        const int kt = blockIdx.y * 4 + kb;
        const long r = r0 + rr;
        if (kt >= KT || r >= Rp) continue;
        const int kk = kb * 16 + 2 * kp;
        const int er = r < R ? e[r] : 0;
        double x0 = ldexp(t[rr][kk], -er), x1 = ldexp(t[rr][kk + 1], -er);
        for (int p = 0; p < s; ++p) {
            const double u0 = x0 * 256.0, u1 = x1 * 256.0;
            const double q0 = trunc(u0), q1 = trunc(u1);
            x0 = u0 - q0;
            x1 = u1 - q1;
            bf16 *o = S + (((long)p * KT + kt) * Rp + r) * 16 + 2 * kp;
            if (f16) {  // the integer itself: 2^-8p is below FP16's range for p >= 3
                o[0] = __builtin_bit_cast(bf16, (_Float16)q0);
                o[1] = __builtin_bit_cast(bf16, (_Float16)q1);
            } else {
                o[0] = (bf16)(float)ldexp(q0, -8 * p);
                o[1] = (bf16)(float)ldexp(q1, -8 * p);
            }
        }
    }
}

// INT8 slices (Ootomo, Ozaki & Yokota, "DGEMM on integer matrix
// multiplication unit", IJHPCA 2024, in the same integer-slice form): the
// scaled entry x (|x| < 1) is cut into s signed 7-bit integers
//   t = 128 x;  q_p = trunc(t);  x = t - q_p   (all exact, |q_p| <= 127)
// so a = 2^(ea-7) sum_p q_p 2^-7p + r with |r| < 2^(ea-7s).  Layout
// S8[p][k/32][row][32] (rows padded to 128, k to 32): a 32-deep k slab of 32
// rows is one contiguous kilobyte, the unit the product kernel moves from
// HBM into LDS.  Same 32 x 64 FP64 tile through LDS as split_kernel; 8
// threads write a row's 32 bytes of a slab, four slices digits per store.
__global__ __launch_bounds__(256) void split_i8_kernel(int R, int Rp, int K, int KT, int s,
                                                       const double *X, long ld, int rcontig,
                                                       const int *e, int8_t *S)
{
    __shared__ double t[32][65];
    const int r0 = blockIdx.x * 32, k0 = blockIdx.y * 64;
    for (int idx = threadIdx.x; idx < 32 * 64; idx += 256) {
        int rr, kk;
        if (rcontig) { rr = idx % 32; kk = idx / 32; }  // r fastest: X[r + k*ld]
        else { kk = idx % 64; rr = idx / 64; }          // k fastest: X[k + r*ld]
        const long r = r0 + rr, k = k0 + kk;
        double v = 0;
        if (r < R && k < K) v = rcontig ? X[r + k * ld] : X[k + r * ld];
        t[rr][kk] = v;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < 2 * 32 * 8; idx += 256) {
        const int kb = idx / 256, rr = (idx % 256) / 8, kg = idx % 8;
        const int kt = blockIdx.y * 2 + kb;
        const long r = r0 + rr;
        if (kt >= KT || r >= Rp) continue;
        const int kk = kb * 32 + 4 * kg;
        const int er = r < R ? e[r] : 0;
        double x[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) x[j] = ldexp(t[rr][kk + j], -er);
        for (int p = 0; p < s; ++p) {
            unsigned w = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const double u = x[j] * 128.0, q = trunc(u);
                x[j] = u - q;
                w |= ((unsigned)(int)q & 0xffu) << (8 * j);
            }
            *(unsigned *)(S + (((long)p * KT + kt) * Rp + r) * 32 + 4 * kg) = w;
        }
    }
}

// ---- slice products ----------------------------------------------------------
typedef const __attribute__((address_space(1))) void *gptr_t;
typedef __attribute__((address_space(3))) void *lptr_t;

// Accumulator set of diagonal d: d = 0, 1, 2, 3 are exact in FP32 over chunks
// of 256, 128, 64 and 64 k (1, 2, 3 and 4 pairs of products below 2^16 each,
// every partial sum a multiple of 2^-8d below 2^(24-8d)); d = 4 and d >= 5
// round, but their terms weigh below 2^-32 of the leading ones, so their FP32
// rounding stays near 2^-56 of the result.  Two rounding sets rather than one
// keep the dependent MFMA chains of a slab short.
__host__ __device__ constexpr int set_of(int d) { return d < 5 ? d : 5; }
__host__ __device__ constexpr int nsets(int s) { return s < 6 ? s : 6; }
// fold period of each set, in 16-deep k slabs
__host__ __device__ constexpr int fold_every(int set)
{
    return set == 0 ? 16 : set == 1 ? 8 : set >= 4 ? 16 : 4;
}

// One workgroup per TM x 64 tile of C, one 32 x 32 MFMA tile per wave (8
// waves on 128 rows, or 4 on 64 rows with two workgroups per CU), so two
// waves share each SIMD.  Per 16-deep k slab all S slices of the A and B
// tiles arrive in LDS by direct HBM->LDS loads through a ring of up to four
// stages (one to three slabs in flight while one is multiplied); each wave
// reads the fragments of every B slice once and runs the pairs p + q <= S - 1
// with v_mfma_f32_32x32x16_bf16, each into the accumulator set of its
// diagonal.  Sets are folded into FP64 on their own period; the epilogue
// applies 2^(ea_i + eb_j - 16), alpha and beta.
// LDS stages of the slab pipeline: as many 16-deep slabs of all S slices as
// fit in the LDS share of one of OCC workgroups, at most four.
__host__ __device__ constexpr int oz_stages(int s, int tm, int occ)
{
    return ((occ == 1 ? 147456 : 163840 / occ) / (s * (tm + 64) * 32)) < 4
               ? ((occ == 1 ? 147456 : 163840 / occ) / (s * (tm + 64) * 32))
               : 4;
}

// Tile of workgroup b.  Workgroups are dispatched round-robin over the 8
// XCDs, each with its own L2: XCD x gets the contiguous range of tile ids
// x*T/8 .. (x+1)*T/8 (when T divides), walked in groups of 8 tile rows, so
// the workgroups one XCD runs together share 8 A row-tiles and a few B
// column-tiles of every k slab instead of streaming them from HBM each.
__device__ __forceinline__ void tile_of(int b, int TM, int TN, int &tm, int &tn)
{
    const int T = TM * TN;
    const int t = (T & 7) == 0 ? (b & 7) * (T >> 3) + (b >> 3) : b;
    constexpr int GM = 8;
    const int per_group = GM * TN;
    const int first = (t / per_group) * GM;
    const int gs = min(TM - first, GM);
    const int r = t % per_group;
    tm = first + r % gs;
    tn = r / gs;
}

typedef _Float16 v8h __attribute__((ext_vector_type(8)));
// one 32x32x16 slice product: BF16 slices are pre-scaled, FP16 slices hold
// the integers (same bits in the LDS stages, read as the MFMA's type)
template <bool F16>
__device__ __forceinline__ v16f slice_mfma(v8bf b, v8bf a, v16f c)
{
    if constexpr (F16)
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(v8h, b),
                                                      __builtin_bit_cast(v8h, a), c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(b, a, c, 0, 0, 0);
}

// one pipeline stage of the slice product: slab kt of every A slice (rows
// 0..TM-1) and B slice (rows TM..TM+TN-1)
template <int NSL, int TM, int TN> struct OzStage {
    __attribute__((aligned(16))) bf16 r[NSL][TM + TN][16];
};

template <int S, int TM, int NW, int OCC, bool F16>
__global__ __launch_bounds__(64 * NW, OCC) void ozaki_gemm_kernel(OzArgs g)
{
    constexpr int NS = nsets(S);
    constexpr int TN = 64;             // tile columns: two columns of waves, 32 each
    constexpr int MW = NW / 2;         // waves down the tile
    constexpr int MI = TM / (32 * MW); // 32-row MFMA tiles per wave
    constexpr int D = oz_stages(S, TM, OCC);
    constexpr int P = (TM + TN) / 32;  // 1 KB pieces of one slice slab
    constexpr int PW = (P + NW - 1) / NW;  // pieces per wave (waves w, w+NW, ...)
    static_assert(!F16 || S <= 6, "FP16 slices: one diagonal per accumulator set");
    static_assert((TM == 64 || TM == 128) && MI >= 1 && D >= 2 && D <= 4 &&
                      PW * S * D < 64,
                  "slab pipeline shape");
    // one variable per stage, so the compiler can tell the slab being read
    // from the slabs still arriving (a shared array would make every LDS read
    // wait for all loads in flight; so would more than a few variables, which
    // is why A and B of a stage share one)
    __shared__ OzStage<S, TM, TN> L0, L1;
    __shared__ OzStage<(D > 2 ? S : 1), TM, TN> L2;
    __shared__ OzStage<(D > 3 ? S : 1), TM, TN> L3;
    auto stage = [&](auto jc) -> bf16(*)[TM + TN][16] {
        constexpr int J = decltype(jc)::value;
        if constexpr (J == 0) return L0.r;
        else if constexpr (J == 1) return L1.r;
        else if constexpr (J == 2) return L2.r;
        else return L3.r;
    };

    int tm, tn;
    tile_of(blockIdx.x, g.tiles_m, g.tiles_n, tm, tn);
    const int m0 = tm * TM, n0 = tn * TN;
    // SYRK: tiles wholly outside the stored triangle have nothing to do
    if (g.tri == 1 && n0 > m0 + TM - 1) return;
    if (g.tri == 2 && n0 + TN - 1 < m0) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave % MW, wn = wave / MW;

    v16f acc[NS][MI];
    double acc64[MI][16];
#pragma unroll
    for (int c = 0; c < NS; ++c)
#pragma unroll
        for (int i = 0; i < MI; ++i) acc[c][i] = v16f{};
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc64[i][r] = 0.0;

    const int KT = g.KT;
    // stage rows 0..TM-1 hold the A tile, TM..TM+TN-1 the B tile; piece
    // x (32 rows, 1 KB per slice, 16 B per lane) is moved by wave x % NW.
    // The two 16-byte halves of stage row r are swapped when bit 3 of r is
    // set, so the 16 lanes of a fragment read hit 64 distinct banks (rows r
    // and r + 8 would share them): a DMA lane writes a fixed LDS slot, so the
    // swap is made on the global side.
    const int half_src = (lane & 1) ^ ((lane >> 4) & 1);
    long src0[PW];
    bool mine[PW], isa[PW];
#pragma unroll
    for (int u = 0; u < PW; ++u) {
        const int x = wave + NW * u;
        mine[u] = x < P;
        isa[u] = 32 * x < TM;
        src0[u] = (isa[u] ? (long)(m0 + 32 * x) : (long)(n0 + 32 * x - TM)) * 16 +
                  (lane >> 1) * 16 + half_src * 8;
    }
    int nmine = 0;
#pragma unroll
    for (int u = 0; u < PW; ++u) nmine += mine[u] ? 1 : 0;
    // this wave's pieces of slice p of slab kt, into stage jc
    auto issue_slice = [&](int kt, int p, auto jc) {
        bf16(*ls)[TM + TN][16] = stage(jc);
#pragma unroll
        for (int u = 0; u < PW; ++u) {
            if (!mine[u]) continue;
            const bf16 *src = isa[u] ? g.SA + ((long)p * KT + kt) * g.Mp * 16 + src0[u]
                                     : g.SB + ((long)p * KT + kt) * g.Np * 16 + src0[u];
            __builtin_amdgcn_global_load_lds((gptr_t)src, (lptr_t)&ls[p][32 * (wave + NW * u)][0],
                                             16, 0, 0);
        }
    };
    auto issue = [&](int kt, auto jc) {
#pragma unroll
        for (int p = 0; p < S; ++p) issue_slice(kt, p, jc);
    };
    auto fold = [&](int c) {
#pragma unroll
        for (int i = 0; i < MI; ++i) {
#pragma unroll
            for (int r = 0; r < 16; ++r)
                acc64[i][r] += F16 ? ldexp((double)acc[c][i][r], -8 * c) : (double)acc[c][i][r];
            acc[c][i] = v16f{};
        }
    };
    // fragments of one slab: every B slice (column block wn) and every A
    // slice (MI row blocks of wm), with the swapped half of each row
    struct Frag {
        v8bf b[S];
        v8bf a[S][MI];
    };
    const int hr = 8 * ((lane >> 5) ^ ((lane >> 3) & 1));
    auto read_b = [&](Frag &f, int q, auto jc) {
        bf16(*ls)[TM + TN][16] = stage(jc);
        f.b[q] = *(const v8bf *)&ls[q][TM + wn * 32 + (lane & 31)][hr];
    };
    auto read_a = [&](Frag &f, int p, auto jc) {
        bf16(*ls)[TM + TN][16] = stage(jc);
#pragma unroll
        for (int i = 0; i < MI; ++i)
            f.a[p][i] = *(const v8bf *)&ls[p][wm * 32 * MI + i * 32 + (lane & 31)][hr];
    };
    // wait until this wave's loads of the slab after the current one are in
    // LDS: at most `later` younger slabs (nmine loads per slice) outstanding
    auto wait_slab = [&](bool drain, auto later_c) {
        constexpr int LATER = decltype(later_c)::value;
        if (drain) __builtin_amdgcn_s_waitcnt(vm_wait(0));
        else if (nmine == PW) __builtin_amdgcn_s_waitcnt(vm_wait(PW * S * LATER));
        else __builtin_amdgcn_s_waitcnt(vm_wait((PW - 1) * S * LATER));
    };
    // A bare barrier after this wave's LDS reads are back: the fence of
    // __syncthreads would also drain the slabs still in flight.
    auto barrier = [&]() {
        __builtin_amdgcn_s_waitcnt(vm_wait(63) & ~(15 << 8));  // lgkmcnt(0)
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_barrier();
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
    };
    Frag cur, nxt;
    // one slab kt, its fragments already in `cur`: wait for slab kt+1,
    // barrier (every wave has its fragments of slab kt, so that stage is
    // free), refill it with slab kt+D, read the fragments of slab kt+1 while
    // multiplying all pairs of slab kt, fold the sets whose chunk ends here
    auto step = [&](int kt, auto jc) {
        constexpr int J = decltype(jc)::value;
        constexpr int J1 = (J + 1) % D;  // stage of slab kt + 1
        const bool more = kt + 1 < KT;
        if (more) {
            wait_slab(kt + D - 1 >= KT, std::integral_constant<int, D - 2>{});
            barrier();
        }
        const bool refill = kt + D < KT;
        // row p of pairs is the last use of A slice p and B slice S-1-p:
        // their registers take the next slab's fragments right after it.
        // Slice p of slab kt+D goes to the freed stage between the rows too,
        // so no wave stalls on a burst of loads before its MFMAs.
#pragma unroll
        for (int p = 0; p < S; ++p) {
#pragma unroll
            for (int q = 0; q + p < S; ++q)
#pragma unroll
                for (int i = 0; i < MI; ++i)
                    acc[set_of(p + q)][i] = slice_mfma<F16>(cur.b[q], cur.a[p][i], acc[set_of(p + q)][i]);
            __builtin_amdgcn_sched_barrier(0);
            if (more) {
                read_a(nxt, p, std::integral_constant<int, J1>{});
                read_b(nxt, S - 1 - p, std::integral_constant<int, J1>{});
            }
            if (refill) issue_slice(kt + D, p, jc);
            __builtin_amdgcn_sched_barrier(0);
        }
        const bool last = !more;
#pragma unroll
        for (int c = 0; c < NS; ++c)
            if (last || (kt + 1) % fold_every(c) == 0) fold(c);
        cur = nxt;
    };

    // all D stages start filled; slab 0's fragments go to registers first
    issue(0, std::integral_constant<int, 0>{});
    if (1 < KT) issue(1, std::integral_constant<int, 1 % D>{});
    if constexpr (D > 2) if (2 < KT) issue(2, std::integral_constant<int, 2 % D>{});
    if constexpr (D > 3) if (3 < KT) issue(3, std::integral_constant<int, 3 % D>{});
    wait_slab(D - 1 >= KT, std::integral_constant<int, D - 1>{});
    barrier();
#pragma unroll
    for (int p = 0; p < S; ++p) {
        read_a(cur, p, std::integral_constant<int, 0>{});
        read_b(cur, p, std::integral_constant<int, 0>{});
    }
    // the first slab is peeled so that the loop head is not the stage-0 step:
    // there the compiler cannot tell stage 0 from the stages still arriving
    // and drains every load in flight
    step(0, std::integral_constant<int, 0>{});
    for (int kt = 1; kt < KT; kt += D) {
        step(kt, std::integral_constant<int, 1 % D>{});
        if constexpr (D > 2) if (kt + 1 < KT) step(kt + 1, std::integral_constant<int, 2 % D>{});
        if constexpr (D > 3) if (kt + 2 < KT) step(kt + 2, std::integral_constant<int, 3 % D>{});
        if (kt + D - 1 < KT) step(kt + D - 1, std::integral_constant<int, 0>{});
    }

    // epilogue: element r of MFMA tile i is C[row][col] with
    //   row = m0 + wm*32*MI + i*32 + lane%32,
    //   col = n0 + wn*32 + 8*(r/4) + 4*(lane/32) + r%4
#pragma unroll
    for (int i = 0; i < MI; ++i) {
        const long gi = m0 + wm * 32 * MI + i * 32 + (lane & 31);
        if (gi >= g.M) continue;
        const int ei = g.ea[gi];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const long gj = n0 + wn * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            if (gj >= g.N) continue;
            if (g.tri == 1 && gi < gj) continue;
            if (g.tri == 2 && gi > gj) continue;
            double *cp = g.C + gi + gj * g.ldc;
            const double v = g.alpha * ldexp(acc64[i][r], ei + g.eb[gj] - 16);
            *cp = g.beta == 0.0 ? v : fma(g.beta, *cp, v);
        }
    }
}

// ---- INT8 slice products ------------------------------------------------------
// v_mfma_i32_32x32x32_i8 sums 32 products of 7-bit integers exactly in
// int32, so every diagonal d = p + q has one exact integer accumulator over
// the whole k range of a launch: diagonal d holds d + 1 <= s pairs, each
// adding at most 32 * 127^2 per 32-deep slab, so a launch covers at most
// i8_max_slabs(s) slabs (k up to 2^31 / (127^2 s), 16k-18k for s = 7, 8;
// longer k runs as several launches accumulating into C).  The epilogue
// weighs diagonal d by 2^-7d in FP64, smallest first, and applies
// 2^(ea_i + eb_j - 14), alpha and beta.
constexpr int i8_max_slabs(int s) { return (int)(2147483647L / ((long)s * 32 * 127 * 127)); }

struct OzI8Args {
    int M, N, Mp, Np, KT;   // KT: 32-deep k slabs stored per slice
    int kt0, nkt;           // slabs [kt0, kt0 + nkt) in this launch
    const int8_t *SA, *SB;  // [s][KT][Mp][32], [s][KT][Np][32]
    const int *ea, *eb;
    double alpha, beta;
    double *C;
    long ldc;
    int tiles_m, tiles_n;
    int tri;
};

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

template <int NSL, int TM, int TN> struct OzStage8 {
    __attribute__((aligned(16))) int8_t r[NSL][TM + TN][32];
};

// Same pipeline as ozaki_gemm_kernel: one workgroup per TM x 64 tile of C,
// one 32 x 32 MFMA tile per wave, 32-deep slabs of all S slices of the A and
// B tiles moved HBM -> LDS by global_load_lds through a ring of up to four
// stages, the next slab's fragments read between the MFMA rows of the
// current one.
template <int S, int TM, int NW, int OCC>
__global__ __launch_bounds__(64 * NW, OCC) void ozaki_i8_kernel(OzI8Args g)
{
    constexpr int TN = 64;
    constexpr int MW = NW / 2;
    constexpr int MI = TM / (32 * MW);
    constexpr int D = oz_stages(S, TM, OCC);
    constexpr int P = (TM + TN) / 32;
    constexpr int PW = (P + NW - 1) / NW;
    static_assert((TM == 64 || TM == 128) && MI >= 1 && D >= 2 && D <= 4 && PW * S * D < 64,
                  "slab pipeline shape");
    __shared__ OzStage8<S, TM, TN> L0, L1;
    __shared__ OzStage8<(D > 2 ? S : 1), TM, TN> L2;
    __shared__ OzStage8<(D > 3 ? S : 1), TM, TN> L3;
    auto stage = [&](auto jc) -> int8_t(*)[TM + TN][32] {
        constexpr int J = decltype(jc)::value;
        if constexpr (J == 0) return L0.r;
        else if constexpr (J == 1) return L1.r;
        else if constexpr (J == 2) return L2.r;
        else return L3.r;
    };

    int tm, tn;
    tile_of(blockIdx.x, g.tiles_m, g.tiles_n, tm, tn);
    const int m0 = tm * TM, n0 = tn * TN;
    if (g.tri == 1 && n0 > m0 + TM - 1) return;
    if (g.tri == 2 && n0 + TN - 1 < m0) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave % MW, wn = wave / MW;

    v16i acc[S][MI];
#pragma unroll
    for (int c = 0; c < S; ++c)
#pragma unroll
        for (int i = 0; i < MI; ++i) acc[c][i] = v16i{};

    const int KT = g.KT, kt0 = g.kt0, NK = g.nkt;
    // piece x (32 rows x 32 B of one slice) is moved by wave x % NW, 16 B
    // per lane; the two halves of a row are swapped when bit 3 of the row is
    // set so that fragment reads are bank-conflict free (as for BF16)
    const int half_src = (lane & 1) ^ ((lane >> 4) & 1);
    long src0[PW];
    bool mine[PW], isa[PW];
#pragma unroll
    for (int u = 0; u < PW; ++u) {
        const int x = wave + NW * u;
        mine[u] = x < P;
        isa[u] = 32 * x < TM;
        src0[u] = (isa[u] ? (long)(m0 + 32 * x) : (long)(n0 + 32 * x - TM)) * 32 +
                  (lane >> 1) * 32 + half_src * 16;
    }
    int nmine = 0;
#pragma unroll
    for (int u = 0; u < PW; ++u) nmine += mine[u] ? 1 : 0;
    auto issue_slice = [&](int kt, int p, auto jc) {
        int8_t(*ls)[TM + TN][32] = stage(jc);
        const long slab = (long)p * KT + kt0 + kt;
#pragma unroll
        for (int u = 0; u < PW; ++u) {
            if (!mine[u]) continue;
            const int8_t *src = isa[u] ? g.SA + slab * g.Mp * 32 + src0[u]
                                       : g.SB + slab * g.Np * 32 + src0[u];
            __builtin_amdgcn_global_load_lds((gptr_t)src, (lptr_t)&ls[p][32 * (wave + NW * u)][0],
                                             16, 0, 0);
        }
    };
    auto issue = [&](int kt, auto jc) {
#pragma unroll
        for (int p = 0; p < S; ++p) issue_slice(kt, p, jc);
    };
    struct Frag {
        v4i b[S];
        v4i a[S][MI];
    };
    const int hr = 16 * ((lane >> 5) ^ ((lane >> 3) & 1));
    auto read_b = [&](Frag &f, int q, auto jc) {
        int8_t(*ls)[TM + TN][32] = stage(jc);
        f.b[q] = *(const v4i *)&ls[q][TM + wn * 32 + (lane & 31)][hr];
    };
    auto read_a = [&](Frag &f, int p, auto jc) {
        int8_t(*ls)[TM + TN][32] = stage(jc);
#pragma unroll
        for (int i = 0; i < MI; ++i)
            f.a[p][i] = *(const v4i *)&ls[p][wm * 32 * MI + i * 32 + (lane & 31)][hr];
    };
    auto wait_slab = [&](bool drain, auto later_c) {
        constexpr int LATER = decltype(later_c)::value;
        if (drain) __builtin_amdgcn_s_waitcnt(vm_wait(0));
        else if (nmine == PW) __builtin_amdgcn_s_waitcnt(vm_wait(PW * S * LATER));
        else __builtin_amdgcn_s_waitcnt(vm_wait((PW - 1) * S * LATER));
    };
    auto barrier = [&]() {
        __builtin_amdgcn_s_waitcnt(vm_wait(63) & ~(15 << 8));  // lgkmcnt(0)
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_barrier();
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
    };
    Frag cur, nxt;
    auto step = [&](int kt, auto jc) {
        constexpr int J = decltype(jc)::value;
        constexpr int J1 = (J + 1) % D;
        const bool more = kt + 1 < NK;
        if (more) {
            wait_slab(kt + D - 1 >= NK, std::integral_constant<int, D - 2>{});
            barrier();
        }
        const bool refill = kt + D < NK;
#pragma unroll
        for (int p = 0; p < S; ++p) {
#pragma unroll
            for (int q = 0; q + p < S; ++q)
#pragma unroll
                for (int i = 0; i < MI; ++i)
                    acc[p + q][i] = __builtin_amdgcn_mfma_i32_32x32x32_i8(cur.b[q], cur.a[p][i],
                                                                          acc[p + q][i], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
            if (more) {
                read_a(nxt, p, std::integral_constant<int, J1>{});
                read_b(nxt, S - 1 - p, std::integral_constant<int, J1>{});
            }
            if (refill) issue_slice(kt + D, p, jc);
            __builtin_amdgcn_sched_barrier(0);
        }
        cur = nxt;
    };

    issue(0, std::integral_constant<int, 0>{});
    if (1 < NK) issue(1, std::integral_constant<int, 1 % D>{});
    if constexpr (D > 2) if (2 < NK) issue(2, std::integral_constant<int, 2 % D>{});
    if constexpr (D > 3) if (3 < NK) issue(3, std::integral_constant<int, 3 % D>{});
    wait_slab(D - 1 >= NK, std::integral_constant<int, D - 1>{});
    barrier();
#pragma unroll
    for (int p = 0; p < S; ++p) {
        read_a(cur, p, std::integral_constant<int, 0>{});
        read_b(cur, p, std::integral_constant<int, 0>{});
    }
    step(0, std::integral_constant<int, 0>{});
    for (int kt = 1; kt < NK; kt += D) {
        step(kt, std::integral_constant<int, 1 % D>{});
        if constexpr (D > 2) if (kt + 1 < NK) step(kt + 1, std::integral_constant<int, 2 % D>{});
        if constexpr (D > 3) if (kt + 2 < NK) step(kt + 2, std::integral_constant<int, 3 % D>{});
        if (kt + D - 1 < NK) step(kt + D - 1, std::integral_constant<int, 0>{});
    }

    // epilogue (the 32 x 32 accumulator layout of the FP32 kernels):
    //   row = m0 + wm*32*MI + i*32 + lane%32,
    //   col = n0 + wn*32 + 8*(r/4) + 4*(lane/32) + r%4
#pragma unroll
    for (int i = 0; i < MI; ++i) {
        const long gi = m0 + wm * 32 * MI + i * 32 + (lane & 31);
        if (gi >= g.M) continue;
        const int ei = g.ea[gi];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const long gj = n0 + wn * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            if (gj >= g.N) continue;
            if (g.tri == 1 && gi < gj) continue;
            if (g.tri == 2 && gi > gj) continue;
            double v = 0.0;
#pragma unroll
            for (int d = S - 1; d >= 0; --d) v += ldexp((double)acc[d][i][r], -7 * d);
            double *cp = g.C + gi + gj * g.ldc;
            v = g.alpha * ldexp(v, ei + g.eb[gj] - 14);
            *cp = g.beta == 0.0 ? v : fma(g.beta, *cp, v);
        }
    }
}

inline bool isN(char c) { return c == 'N' || c == 'n'; }

struct Split {
    bf16 *S;      // BF16 / FP16 slices
    int8_t *S8;   // INT8 slices
    int *e;
    int Rp;
};

// K chunks of the strided exponent pass: enough for about 2048 workgroups,
// each at least 64 deep
int exp_chunks(int R, int K)
{
    return std::max(1, std::min(ceil_div(2048, ceil_div(R, 256)), K / 64));
}

// Workspace of one split call: the non-finite flag, then room for the
// partial line maxima of the strided exponent pass.
struct SplitWs {
    Scratch scr;
    int *bad = nullptr;
    double *part = nullptr;
};
SplitWs split_workspace(hipStream_t st, int R1, int R2, int K)
{
    const size_t part = (size_t)std::max((long)exp_chunks(R1, K) * R1, (long)exp_chunks(R2, K) * R2);
    SplitWs ws;
    ws.scr = dev_scratch(256 + part * sizeof(double), st);
    char *w = (char *)ws.scr.p;
    if (!w) return ws;
    ws.bad = (int *)w;
    ws.part = (double *)(w + 256);
    return ws;
}

// Exponents and slices of an R x K operand given by (X, ld, rcontig); KT is
// the number of k slabs (16 deep for BF16 / FP16, 32 deep for INT8).
int make_split(hipStream_t st, int R, int K, int KT, int s, const double *X, long ld, int rcontig,
               int fmt, Split &out, int *bad, double *part)
{
    const int f16 = fmt == GPUBLAS_SLICE_FP16, i8 = fmt == GPUBLAS_SLICE_INT8;
    out.Rp = ceil_div(R, 128) * 128;
    out.e = (int *)gpublas_dev_alloc((size_t)R * sizeof(int));
    // both formats: 32 bytes per row of a slab
    void *sl = gpublas_dev_alloc((size_t)s * KT * out.Rp * 32);
    out.S = i8 ? nullptr : (bf16 *)sl;
    out.S8 = i8 ? (int8_t *)sl : nullptr;
    if (!out.e || !sl) return (int)hipErrorOutOfMemory;
    PROF_BEGIN(st, "ozaki_split");
    if (rcontig) {
        const int chunks = exp_chunks(R, K);
        hipLaunchKernelGGL(exp_rows_strided, dim3((unsigned)ceil_div(R, 256), (unsigned)chunks),
                           dim3(256), 0, st, R, K, X, ld, part, bad);
        hipLaunchKernelGGL(exp_rows_finish, dim3((unsigned)ceil_div(R, 256)), dim3(256), 0, st, R,
                           chunks, part, out.e);
    } else
        hipLaunchKernelGGL(exp_rows_contig, dim3((unsigned)ceil_div(R, 4)), dim3(256), 0, st, R, K,
                           X, ld, out.e, bad);
    if (i8)
        hipLaunchKernelGGL(split_i8_kernel, dim3((unsigned)(out.Rp / 32), (unsigned)ceil_div(KT, 2)),
                           dim3(256), 0, st, R, out.Rp, K, KT, s, X, ld, rcontig, out.e, out.S8);
    else
        hipLaunchKernelGGL(split_kernel, dim3((unsigned)(out.Rp / 32), (unsigned)ceil_div(KT, 4)),
                           dim3(256), 0, st, R, out.Rp, K, KT, s, X, ld, rcontig, out.e, f16,
                           out.S);
    PROF_END(st);
    return (int)hipGetLastError();
}

void free_split(Split &x)
{
    gpublas_dev_release(x.S ? (void *)x.S : (void *)x.S8);
    gpublas_dev_release(x.e);
}

template <int S>
void launch_i8(hipStream_t st, unsigned blocks, const OzI8Args &g)
{
    constexpr int TM = oz_tile_rows(S), NW = TM / 16, OCC = 8 / NW;
    hipLaunchKernelGGL((ozaki_i8_kernel<S, TM, NW, OCC>), dim3(blocks), dim3(64 * NW), 0, st, g);
}

// INT8 slice products over all KT slabs: one launch per i8_max_slabs(s)
// slabs, the later ones accumulating into C (beta 1).
int run_products_i8(hipStream_t st, int M, int N, int KT, int s, const Split &a, const Split &b,
                    double alpha, double beta, double *C, int ldc, int tri)
{
    OzI8Args g;
    g.M = M;
    g.N = N;
    g.Mp = a.Rp;
    g.Np = b.Rp;
    g.KT = KT;
    g.SA = a.S8;
    g.SB = b.S8;
    g.ea = a.e;
    g.eb = b.e;
    g.alpha = alpha;
    g.C = C;
    g.ldc = ldc;
    g.tri = tri;
    g.tiles_m = ceil_div(M, oz_tile_rows(s));
    g.tiles_n = ceil_div(N, 64);
    const unsigned blocks = (unsigned)((long)g.tiles_m * g.tiles_n);
    const int per = i8_max_slabs(s);
    PROF_BEGIN(st, "ozaki_i8");
    for (int kt0 = 0; kt0 < KT; kt0 += per) {
        g.kt0 = kt0;
        g.nkt = std::min(per, KT - kt0);
        g.beta = kt0 == 0 ? beta : 1.0;
        switch (s) {
        case 1: launch_i8<1>(st, blocks, g); break;
        case 2: launch_i8<2>(st, blocks, g); break;
        case 3: launch_i8<3>(st, blocks, g); break;
        case 4: launch_i8<4>(st, blocks, g); break;
        case 5: launch_i8<5>(st, blocks, g); break;
        case 6: launch_i8<6>(st, blocks, g); break;
        case 7: launch_i8<7>(st, blocks, g); break;
        default: launch_i8<8>(st, blocks, g); break;
        }
    }
    PROF_END(st);
    return (int)hipGetLastError();
}

template <int S, bool F16>
void launch_products(hipStream_t st, unsigned blocks, const OzArgs &g)
{
    constexpr int TM = oz_tile_rows(S), NW = TM / 16, OCC = 8 / NW;
    hipLaunchKernelGGL((ozaki_gemm_kernel<S, TM, NW, OCC, F16>), dim3(blocks), dim3(64 * NW), 0, st,
                       g);
}

int run_products(hipStream_t st, int M, int N, int KT, int s, int f16, const Split &a,
                 const Split &b, double alpha, double beta, double *C, int ldc, int tri)
{
    OzArgs g;
    g.M = M;
    g.N = N;
    g.Mp = a.Rp;
    g.Np = b.Rp;
    g.KT = KT;
    g.SA = a.S;
    g.SB = b.S;
    g.ea = a.e;
    g.eb = b.e;
    g.alpha = alpha;
    g.beta = beta;
    g.C = C;
    g.ldc = ldc;
    g.tri = tri;
    g.tiles_m = ceil_div(M, oz_tile_rows(s));
    g.tiles_n = ceil_div(N, 64);
    const unsigned blocks = (unsigned)((long)g.tiles_m * g.tiles_n);
    PROF_BEGIN(st, "ozaki_gemm");
    if (f16) {
        switch (s) {
        case 1: launch_products<1, true>(st, blocks, g); break;
        case 2: launch_products<2, true>(st, blocks, g); break;
        case 3: launch_products<3, true>(st, blocks, g); break;
        case 4: launch_products<4, true>(st, blocks, g); break;
        case 5: launch_products<5, true>(st, blocks, g); break;
        default: launch_products<6, true>(st, blocks, g); break;
        }
    } else {
        switch (s) {
        case 1: launch_products<1, false>(st, blocks, g); break;
        case 2: launch_products<2, false>(st, blocks, g); break;
        case 3: launch_products<3, false>(st, blocks, g); break;
        case 4: launch_products<4, false>(st, blocks, g); break;
        case 5: launch_products<5, false>(st, blocks, g); break;
        case 6: launch_products<6, false>(st, blocks, g); break;
        case 7: launch_products<7, false>(st, blocks, g); break;
        default: launch_products<8, false>(st, blocks, g); break;
        }
    }
    PROF_END(st);
    return (int)hipGetLastError();
}

}  // namespace

int gpublas_gemm_d(hipStream_t s, char ta, char tb, int m, int n, int k, double alpha,
                   const double *A, int lda, const double *B, int ldb, double beta, double *C,
                   int ldc);
int gpublas_dev_dsyrk(void *stream, char uplo, char trans, int n, int k, double alpha,
                      const double *d_A, int lda, double beta, double *d_C, int ldc);

extern "C" {

int gpublas_dev_dgemm_split_fmt(void *stream, int slices, int fmt, char transa, char transb, int m,
                                int n, int k, double alpha, const double *d_A, int lda,
                                const double *d_B, int ldb, double beta, double *d_C, int ldc)
{
    hipStream_t st = pick_stream(stream);
    if (m <= 0 || n <= 0) return 0;
    if (fmt != GPUBLAS_SLICE_BF16 && fmt != GPUBLAS_SLICE_FP16 && fmt != GPUBLAS_SLICE_INT8)
        return (int)hipErrorInvalidValue;
    const int maxs = fmt == GPUBLAS_SLICE_FP16 ? 6 : OZ_MAXS;
    if (slices < 1 || slices > maxs) return (int)hipErrorInvalidValue;
    if (alpha == 0.0 || k <= 0)
        return gpublas_gemm_d(st, transa, transb, m, n, k, alpha, d_A, lda, d_B, ldb, beta, d_C, ldc);
    const int i8 = fmt == GPUBLAS_SLICE_INT8;
    const int KT = ceil_div(k, i8 ? 32 : 16);
    SplitWs ws = split_workspace(st, m, n, k);
    if (!ws.bad) return (int)hipErrorOutOfMemory;
    int *bad = ws.bad;
    (void)hipMemsetAsync(bad, 0, sizeof(int), st);
    Split a{}, b{};
    // op(A) rows: A 'N' -> X[r + k*lda] (rows contiguous); 'T' -> X[k + r*lda]
    int err = make_split(st, m, k, KT, slices, d_A, lda, isN(transa), fmt, a, bad, ws.part);
    // op(B) columns: B 'N' -> element (k, j) at B[k + j*ldb] (k contiguous); 'T' -> B[j + k*ldb]
    if (!err) err = make_split(st, n, k, KT, slices, d_B, ldb, !isN(transb), fmt, b, bad, ws.part);
    int hbad = 0;
    if (!err) err = (int)hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st);
    if (!err) err = (int)hipStreamSynchronize(st);
    if (!err) {
        if (hbad)  // Inf / NaN input: the splitting is undefined, use the FP64 kernel
            err = gpublas_gemm_d(st, transa, transb, m, n, k, alpha, d_A, lda, d_B, ldb, beta, d_C, ldc);
        else if (i8)
            err = run_products_i8(st, m, n, KT, slices, a, b, alpha, beta, d_C, ldc, 0);
        else
            err = run_products(st, m, n, KT, slices, fmt == GPUBLAS_SLICE_FP16, a, b, alpha, beta,
                               d_C, ldc, 0);
    }
    (void)hipStreamSynchronize(st);
    free_split(a);
    free_split(b);
    return err;
}

int gpublas_dev_dgemm_split(void *stream, int slices, char transa, char transb, int m, int n,
                            int k, double alpha, const double *d_A, int lda, const double *d_B,
                            int ldb, double beta, double *d_C, int ldc)
{
    return gpublas_dev_dgemm_split_fmt(stream, slices, GPUBLAS_SLICE_BF16, transa, transb, m, n, k,
                                       alpha, d_A, lda, d_B, ldb, beta, d_C, ldc);
}

int gpublas_dev_dsyrk_split_fmt(void *stream, int slices, int fmt, char uplo, char trans, int n,
                                int k, double alpha, const double *d_A, int lda, double beta,
                                double *d_C, int ldc)
{
    hipStream_t st = pick_stream(stream);
    if (n <= 0) return 0;
    if (fmt != GPUBLAS_SLICE_BF16 && fmt != GPUBLAS_SLICE_FP16 && fmt != GPUBLAS_SLICE_INT8)
        return (int)hipErrorInvalidValue;
    const int maxs = fmt == GPUBLAS_SLICE_FP16 ? 6 : OZ_MAXS;
    if (slices < 1 || slices > maxs) return (int)hipErrorInvalidValue;
    if (alpha == 0.0 || k <= 0)
        return gpublas_dev_dsyrk(st, uplo, trans, n, k, alpha, d_A, lda, beta, d_C, ldc);
    const int i8 = fmt == GPUBLAS_SLICE_INT8;
    const int KT = ceil_div(k, i8 ? 32 : 16);
    SplitWs ws = split_workspace(st, n, n, k);
    if (!ws.bad) return (int)hipErrorOutOfMemory;
    int *bad = ws.bad;
    (void)hipMemsetAsync(bad, 0, sizeof(int), st);
    Split a{};
    // op(A) = A (trans N: rows contiguous) or A^T; op(B) = op(A)^T has the same lines
    int err = make_split(st, n, k, KT, slices, d_A, lda, isN(trans), fmt, a, bad, ws.part);
    int hbad = 0;
    if (!err) err = (int)hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st);
    if (!err) err = (int)hipStreamSynchronize(st);
    if (!err) {
        const int tri = (uplo == 'U' || uplo == 'u') ? 2 : 1;
        if (hbad) err = gpublas_dev_dsyrk(st, uplo, trans, n, k, alpha, d_A, lda, beta, d_C, ldc);
        else if (i8) err = run_products_i8(st, n, n, KT, slices, a, a, alpha, beta, d_C, ldc, tri);
        else err = run_products(st, n, n, KT, slices, fmt == GPUBLAS_SLICE_FP16, a, a, alpha, beta,
                                d_C, ldc, tri);
    }
    (void)hipStreamSynchronize(st);
    free_split(a);
    return err;
}

int gpublas_dev_dsyrk_split(void *stream, int slices, char uplo, char trans, int n, int k,
                            double alpha, const double *d_A, int lda, double beta, double *d_C,
                            int ldc)
{
    return gpublas_dev_dsyrk_split_fmt(stream, slices, GPUBLAS_SLICE_BF16, uplo, trans, n, k, alpha,
                                       d_A, lda, beta, d_C, ldc);
}

}  // extern "C"
