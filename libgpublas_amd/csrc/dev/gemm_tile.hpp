// gemm_tile.hpp -- the generic LDS-tiled MFMA GEMM tile shared by the
// per-call kernels (gemm.hip) and the grouped persistent kernels of the
// serving loop (group.hip).  Both paths compute a tile of C with the same
// code, so a product's bits do not depend on how it was launched.
//
// Tile scheme:
//  * a 256-thread workgroup owns a BM x BN tile of C; its 4 waves split the
//    tile 2 x 2 and each wave accumulates (BM/2) x (BN/2) with
//    v_mfma_f64_16x16x4_f64 (FP64) or v_mfma_f32_16x16x4_f32 (FP32);
//  * op(A) and op(B) tiles of depth BK stream through double-buffered LDS,
//    stored k-major ([BK][BM+PAD]) so each MFMA operand fragment is one
//    ds_read per lane; the next k-tile's global loads are in flight while
//    the current one is multiplied;
//  * the MFMA is issued with the B fragment first, so the accumulator holds
//    C^T and 16 consecutive lanes own 16 consecutive rows of one column of
//    C: the epilogue reads/writes C in 128-byte column segments;
//  * the k index advances in steps of 4 in increasing order whatever BM, BN
//    and BK are, so every tile shape gives the same result bits.
#pragma once
#include "devcommon.hpp"
#include <math.h>

namespace {

enum { PAT_OUTER = 0, PAT_KCONT = 1, PAT_ELEM = 2 };
// Template-only patterns of the fast tile: full tiles of an operand that is
// not 16-B aligned (or whose leading dimension is odd), read in the OUTER /
// KCONT layout with one element per load and no bounds checks.  Same values
// in the same LDS places as the aligned patterns, hence the same bits.
enum { PAT_OUTER_U = 3, PAT_KCONT_U = 4 };
__host__ __device__ constexpr int pat_base(int p) { return p >= PAT_OUTER_U ? p - PAT_OUTER_U : p; }
__host__ __device__ constexpr bool pat_unaligned(int p) { return p >= PAT_OUTER_U; }
enum { SYM_NONE = 0, SYM_LOWER = 1, SYM_UPPER = 2 };

// One operand of the product as the loader sees it: element (o, k) of the
// operand tile, where o is the row of op(A) (or the column of op(B)) and k
// the reduction index.  `form` says which of the two address forms is the
// natural one: OUTER -> p[o + k*ld], KCONT -> p[k + o*ld].  For a symmetric
// operand (sym != 0) the stored matrix S is addressed S[r + c*ld] with
// (r, c) = (o, k) for OUTER and (k, o) for KCONT, and only the `sym`
// triangle of S may be read.
template <typename T>
struct Operand {
    const T *p;
    long ld;
    int olim;   // rows of op(A) / columns of op(B)
    int form;   // PAT_OUTER or PAT_KCONT
    int sym;    // SYM_*
    int vec_ok; // p 16-B aligned and ld a multiple of the vector width
};

template <typename T>
struct GemmArgs {
    int m, n, k;
    T alpha, beta;
    Operand<T> a, b;
    T *C;
    long ldc;
    int tiles_m, tiles_n;
    int tri;  // 0 rectangular grid, SYM_LOWER / SYM_UPPER: triangular tile grid + mask
    int full_m, full_n;  // generic kernel: tiles below both counts were done by the fast kernel
};

__device__ __forceinline__ bool stored(int sym, long r, long c)
{
    return sym == SYM_LOWER ? r >= c : r <= c;
}

// Pattern for the tile o in [o0, o0+R), k in [k0, k0+BK).
// LDS tile layout: element (k, o) of a staged tile lives at column
// o ^ lds_swz<T>(k).  Rows are BM + 128 bytes apart (0 modulo the 64 banks
// for every other row), so the transposed stores of a k-contiguous operand
// would put the 8 (double) or 4 (float) k rows one wave writes for a
// column on the same banks; the XOR moves each row pair (quad) to its own
// bank group.  It keeps 16-byte vectors and 32-column blocks intact, so
// vector stores and the MFMA fragment reads stay conflict-free.
template <typename T> __device__ __forceinline__ int lds_swz(int k)
{
    return sizeof(T) == 8 ? 4 * (k >> 1) : 16 * (k >> 2);
}

template <typename T, int R, int BK>
__device__ __forceinline__ int tile_pattern(const Operand<T> &op, int o0, int k0, int K)
{
    const bool full = (o0 + R <= op.olim) && (k0 + BK <= K);
    if (op.sym == SYM_NONE) return (full && op.vec_ok) ? op.form : PAT_ELEM;
    if (!full || !op.vec_ok) return PAT_ELEM;
    // (r, c) ranges of the tile in the stored matrix
    long rlo, rhi, clo, chi;
    if (op.form == PAT_OUTER) { rlo = o0; rhi = o0 + R - 1; clo = k0; chi = k0 + BK - 1; }
    else { rlo = k0; rhi = k0 + BK - 1; clo = o0; chi = o0 + R - 1; }
    bool all_direct, all_mirror;
    if (op.sym == SYM_LOWER) { all_direct = rlo >= chi; all_mirror = rhi <= clo; }
    else { all_direct = rhi <= clo; all_mirror = rlo >= chi; }
    if (all_direct) return op.form;
    if (all_mirror) return op.form == PAT_OUTER ? PAT_KCONT : PAT_OUTER;
    return PAT_ELEM;
}

template <typename T>
__device__ __forceinline__ T elem_load(const Operand<T> &op, long o, long k, int K)
{
    if (o >= op.olim || k >= K) return T(0);
    long r = op.form == PAT_OUTER ? o : k;
    long c = op.form == PAT_OUTER ? k : o;
    if (op.sym != SYM_NONE && !stored(op.sym, r, c)) { long t = r; r = c; c = t; }
    return op.p[r + c * op.ld];
}

// Global -> registers.  Each thread owns NV vectors of VEC elements.  For
// PAT_OUTER and PAT_ELEM the vector runs along o; for PAT_KCONT along k.
template <typename T, int R, int BK>
struct TileLoader {
    static constexpr int VEC = 16 / (int)sizeof(T);
    static constexpr int NV = R * BK / VEC / 256;
    static_assert(NV >= 1, "tile too small for 256 threads");
    static_assert((R & (R - 1)) == 0 &&
                      (sizeof(T) == 8 ? 4 * ((BK - 1) >> 1) : 16 * ((BK - 1) >> 2)) < R,
                  "lds_swz must stay inside the tile rows");
    typedef T Vec __attribute__((ext_vector_type(VEC)));

    __device__ __forceinline__ static void load(const Operand<T> &op, int o0, int k0, int K,
                                                int pat, Vec (&st)[NV])
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int r = 0; r < NV; ++r) {
            const int v = t + r * 256;
            if (pat == PAT_OUTER) {
                const int o = (v % (R / VEC)) * VEC, kk = v / (R / VEC);
                st[r] = *(const Vec *)(op.p + (long)(o0 + o) + (long)(k0 + kk) * op.ld);
            } else if (pat == PAT_KCONT) {
                const int kk = (v % (BK / VEC)) * VEC, o = v / (BK / VEC);
                st[r] = *(const Vec *)(op.p + (long)(k0 + kk) + (long)(o0 + o) * op.ld);
            } else {
                const int o = (v % (R / VEC)) * VEC, kk = v / (R / VEC);
#pragma unroll
                for (int e = 0; e < VEC; ++e) st[r][e] = elem_load(op, o0 + o + e, k0 + kk, K);
            }
        }
    }

    // Fast path: full tile, pattern fixed at compile time; a vector load per
    // VEC elements when the operand is aligned (UA false), one load per
    // element otherwise.
    template <int PAT, bool UA = false>
    __device__ __forceinline__ static void load_pat(const Operand<T> &op, int o0, int k0,
                                                    Vec (&st)[NV])
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int r = 0; r < NV; ++r) {
            const int v = t + r * 256;
            if (PAT == PAT_OUTER) {
                const int o = (v % (R / VEC)) * VEC, kk = v / (R / VEC);
                const T *src = op.p + (long)(o0 + o) + (long)(k0 + kk) * op.ld;
                if (UA) {
#pragma unroll
                    for (int e = 0; e < VEC; ++e) st[r][e] = src[e];
                } else {
                    st[r] = *(const Vec *)src;
                }
            } else {
                const int kk = (v % (BK / VEC)) * VEC, o = v / (BK / VEC);
                const T *src = op.p + (long)(k0 + kk) + (long)(o0 + o) * op.ld;
                if (UA) {
#pragma unroll
                    for (int e = 0; e < VEC; ++e) st[r][e] = src[e];
                } else {
                    st[r] = *(const Vec *)src;
                }
            }
        }
    }

    // A ragged tile with the pattern fixed at compile time: element loads,
    // zero outside the operand (o >= olim or k >= K; the address is clamped
    // to the operand's first element there, so no load is skipped by a
    // branch), vectors laid out as load_pat's.
    template <int PAT>
    __device__ __forceinline__ static void load_guard(const Operand<T> &op, int o0, int k0, int K,
                                                      Vec (&st)[NV])
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int r = 0; r < NV; ++r) {
            const int v = t + r * 256;
            const int o = PAT == PAT_OUTER ? (v % (R / VEC)) * VEC : v / (BK / VEC);
            const int kk = PAT == PAT_OUTER ? v / (R / VEC) : (v % (BK / VEC)) * VEC;
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                const long oe = o0 + o + (PAT == PAT_OUTER ? e : 0);
                const long ke = k0 + kk + (PAT == PAT_OUTER ? 0 : e);
                const bool in = oe < op.olim && ke < K;
                const long off = PAT == PAT_OUTER ? oe + ke * op.ld : ke + oe * op.ld;
                const T x = op.p[in ? off : 0];
                st[r][e] = in ? x : T(0);
            }
        }
    }

    template <int PAT, int LDS_LD>
    __device__ __forceinline__ static void store_pat(T (*s)[LDS_LD], const Vec (&st)[NV])
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int r = 0; r < NV; ++r) {
            const int v = t + r * 256;
            if (PAT == PAT_KCONT) {
                const int kk = (v % (BK / VEC)) * VEC, o = v / (BK / VEC);
#pragma unroll
                for (int e = 0; e < VEC; ++e) s[kk + e][o ^ lds_swz<T>(kk + e)] = st[r][e];
            } else {
                const int o = (v % (R / VEC)) * VEC, kk = v / (R / VEC);
                *(Vec *)&s[kk][o ^ lds_swz<T>(kk)] = st[r];
            }
        }
    }

    template <int LDS_LD>
    __device__ __forceinline__ static void store(T (*s)[LDS_LD], int pat, const Vec (&st)[NV])
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int r = 0; r < NV; ++r) {
            const int v = t + r * 256;
            if (pat == PAT_KCONT) {
                const int kk = (v % (BK / VEC)) * VEC, o = v / (BK / VEC);
#pragma unroll
                for (int e = 0; e < VEC; ++e) s[kk + e][o ^ lds_swz<T>(kk + e)] = st[r][e];
            } else {
                const int o = (v % (R / VEC)) * VEC, kk = v / (R / VEC);
                *(Vec *)&s[kk][o ^ lds_swz<T>(kk)] = st[r];
            }
        }
    }
};

template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef v4d Acc;
    __device__ __forceinline__ static Acc mma(double x, double y, Acc c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
    }
    // accumulator element r of lane l holds D[row][col] with
    // row = (l/16) + 4r, col = l%16 (D = C^T tile)
    __device__ __forceinline__ static int drow(int l, int r) { return (l >> 4) + 4 * r; }
};
template <> struct Mfma<float> {
    typedef v4f Acc;
    __device__ __forceinline__ static Acc mma(float x, float y, Acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, c, 0, 0, 0);
    }
    __device__ __forceinline__ static int drow(int l, int r) { return 4 * (l >> 4) + r; }
};

__device__ __forceinline__ void tri_decode(int b, int &hi, int &lo)
{
    // b-th pair (hi, lo) with hi >= lo in row-major order of a lower triangle
    int h = (int)((sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
    while (h * (h + 1) / 2 > b) --h;
    while ((h + 1) * (h + 2) / 2 <= b) ++h;
    hi = h;
    lo = b - h * (h + 1) / 2;
}

// C tile (tm, tn) of g: C = alpha*op(A)*op(B) + beta*C restricted to the
// tile (and to the triangle g.tri).  As / Bs: LDS staging of the caller.
// Ends with all waves past their last LDS read, so the caller may start
// another tile at once.
template <typename T, int BM, int BN, int BK>
__device__ __forceinline__ void gemm_tile(const GemmArgs<T> &g, int tm, int tn,
                                          T (*As)[BK][BM + 128 / sizeof(T)],
                                          T (*Bs)[BK][BN + 128 / sizeof(T)])
{
    constexpr int PAD = 128 / (int)sizeof(T);
    constexpr int WM = BM / 2, WN = BN / 2, FM = WM / 16, FN = WN / 16;
    typedef TileLoader<T, BM, BK> LA;
    typedef TileLoader<T, BN, BK> LB;
    typedef typename Mfma<T>::Acc Acc;

    const int m0 = tm * BM, n0 = tn * BN;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    Acc acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = Acc{};

    typename LA::Vec sa[LA::NV];
    typename LB::Vec sb[LB::NV];
    const int K = g.k;
    const int ktiles = (K + BK - 1) / BK;

    int pa = tile_pattern<T, BM, BK>(g.a, m0, 0, K);
    int pb = tile_pattern<T, BN, BK>(g.b, n0, 0, K);
    LA::load(g.a, m0, 0, K, pa, sa);
    LB::load(g.b, n0, 0, K, pb, sb);
    LA::template store<BM + PAD>(As[0], pa, sa);
    LB::template store<BN + PAD>(Bs[0], pb, sb);
    __syncthreads();

    for (int kt = 0; kt < ktiles; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < ktiles;
        if (more) {
            const int k1 = (kt + 1) * BK;
            pa = tile_pattern<T, BM, BK>(g.a, m0, k1, K);
            pb = tile_pattern<T, BN, BK>(g.b, n0, k1, K);
            LA::load(g.a, m0, k1, K, pa, sa);
            LB::load(g.b, n0, k1, K, pb, sb);
        }
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            T fa[FM], fb[FN];
            const int kr = kk + (lane >> 4);
            // (c0 + l) ^ swz = (c0 ^ shi) + (l ^ slo) for c0 a multiple of 16 and
            // l < 16: the high part of the swizzle depends on kk alone (a
            // constant of the unrolled loop), so the fragment offsets stay
            // immediate; the low part is one lane-dependent term per step
            const int shi = lds_swz<T>(kk) & ~15;
            const int lo = (lane & 15) ^ (lds_swz<T>(kr) & 15);
#pragma unroll
            for (int i = 0; i < FM; ++i)
                fa[i] = As[cur][kr][((wm * WM + i * 16) ^ shi) + lo];
#pragma unroll
            for (int j = 0; j < FN; ++j)
                fb[j] = Bs[cur][kr][((wn * WN + j * 16) ^ shi) + lo];
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) acc[i][j] = Mfma<T>::mma(fb[j], fa[i], acc[i][j]);
        }
        if (more) {
            LA::template store<BM + PAD>(As[cur ^ 1], pa, sa);
            LB::template store<BN + PAD>(Bs[cur ^ 1], pb, sb);
        }
        __syncthreads();
    }

    // Epilogue: accumulator (i, j) element r of this lane is C[row][col] with
    // row = m0 + wm*WM + i*16 + lane%16, col = n0 + wn*WN + j*16 + drow(lane, r).
    const T alpha = g.alpha, beta = g.beta;
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        const long gi = m0 + wm * WM + i * 16 + (lane & 15);
        if (gi >= g.m) continue;
#pragma unroll
        for (int j = 0; j < FN; ++j) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const long gj = n0 + wn * WN + j * 16 + Mfma<T>::drow(lane, r);
                if (gj >= g.n) continue;
                if (g.tri == SYM_LOWER && gi < gj) continue;
                if (g.tri == SYM_UPPER && gi > gj) continue;
                T *c = g.C + gi + gj * g.ldc;
                T v = alpha * acc[i][j][r];
                if (beta != T(0)) v += beta * *c;
                *c = v;
            }
        }
    }
}

// Full C tile (tm, tn) of g with access patterns PA / PB fixed at compile
// time: no per-element selects or bounds checks in the main loop (only a
// partial last k-tile takes guarded element loads).  PAT_OUTER / PAT_KCONT:
// both operands 16-B aligned with leading dimensions a multiple of the
// vector width (vector loads); PAT_OUTER_U / PAT_KCONT_U: any alignment (one
// load per element).  Same k order and epilogue as gemm_tile, hence the same
// bits.
// EDGE: a ragged tile (rows past m or columns past n, any k): every load
// guarded (load_guard) and the epilogue bounded; elements past an operand's
// end read as zero, as in gemm_tile, so the bits are still the same.
template <typename T, int BM, int BN, int BK, int PA_, int PB_, int TRI, bool EDGE = false>
__device__ __forceinline__ void gemm_fast_tile(const GemmArgs<T> &g, int tm, int tn,
                                               T (*As)[BK][BM + 128 / sizeof(T)],
                                               T (*Bs)[BK][BN + 128 / sizeof(T)])
{
    constexpr int PA = pat_base(PA_), PB = pat_base(PB_);
    constexpr bool UA = pat_unaligned(PA_), UB = pat_unaligned(PB_);
    constexpr int PAD = 128 / (int)sizeof(T);
    constexpr int WM = BM / 2, WN = BN / 2, FM = WM / 16, FN = WN / 16;
    typedef TileLoader<T, BM, BK> LA;
    typedef TileLoader<T, BN, BK> LB;
    typedef typename Mfma<T>::Acc Acc;

    const int m0 = tm * BM, n0 = tn * BN;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int K = g.k;
    const int ktiles = (K + BK - 1) / BK;

    Acc acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = Acc{};

    typename LA::Vec sa[LA::NV];
    typename LB::Vec sb[LB::NV];
    auto load = [&](int k0) -> bool {
        if (EDGE) {
            LA::template load_guard<PA>(g.a, m0, k0, K, sa);
            LB::template load_guard<PB>(g.b, n0, k0, K, sb);
            return false;  // laid out as the patterns
        }
        if (k0 + BK <= K) {
            LA::template load_pat<PA, UA>(g.a, m0, k0, sa);
            LB::template load_pat<PB, UB>(g.b, n0, k0, sb);
            return false;
        }
        LA::load(g.a, m0, k0, K, PAT_ELEM, sa);
        LB::load(g.b, n0, k0, K, PAT_ELEM, sb);
        return true;
    };
    auto store = [&](int buf, bool tail) {
        if (!tail) {
            LA::template store_pat<PA, BM + PAD>(As[buf], sa);
            LB::template store_pat<PB, BN + PAD>(Bs[buf], sb);
        } else {
            LA::template store_pat<PAT_OUTER, BM + PAD>(As[buf], sa);
            LB::template store_pat<PAT_OUTER, BN + PAD>(Bs[buf], sb);
        }
    };

    store(0, load(0));
    __syncthreads();
    for (int kt = 0; kt < ktiles; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < ktiles;
        bool tail = false;
        if (more) tail = load((kt + 1) * BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            T fa[FM], fb[FN];
            const int kr = kk + (lane >> 4);
            // (c0 + l) ^ swz = (c0 ^ shi) + (l ^ slo) for c0 a multiple of 16 and
            // l < 16: the high part of the swizzle depends on kk alone (a
            // constant of the unrolled loop), so the fragment offsets stay
            // immediate; the low part is one lane-dependent term per step
            const int shi = lds_swz<T>(kk) & ~15;
            const int lo = (lane & 15) ^ (lds_swz<T>(kr) & 15);
#pragma unroll
            for (int i = 0; i < FM; ++i)
                fa[i] = As[cur][kr][((wm * WM + i * 16) ^ shi) + lo];
#pragma unroll
            for (int j = 0; j < FN; ++j)
                fb[j] = Bs[cur][kr][((wn * WN + j * 16) ^ shi) + lo];
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) acc[i][j] = Mfma<T>::mma(fb[j], fa[i], acc[i][j]);
        }
        if (more) store(cur ^ 1, tail);
        __syncthreads();
    }

    const T alpha = g.alpha, beta = g.beta;
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        const long gi = m0 + wm * WM + i * 16 + (lane & 15);
#pragma unroll
        for (int j = 0; j < FN; ++j) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const long gj = n0 + wn * WN + j * 16 + Mfma<T>::drow(lane, r);
                if (EDGE && (gi >= g.m || gj >= g.n)) continue;
                if (TRI == SYM_LOWER && gi < gj) continue;
                if (TRI == SYM_UPPER && gi > gj) continue;
                T *c = g.C + gi + gj * g.ldc;
                const T v = alpha * acc[i][j][r];
                *c = beta == T(0) ? v : v + beta * *c;
            }
        }
    }
}

inline bool is_n(char c) { return c == 'N' || c == 'n'; }
inline bool is_l(char c) { return c == 'L' || c == 'l'; }
inline bool is_u(char c) { return c == 'U' || c == 'u'; }

template <typename T>
Operand<T> make_op(const T *p, long ld, int olim, int form, int sym)
{
    Operand<T> o;
    o.p = p;
    o.ld = ld;
    o.olim = olim;
    o.form = form;
    o.sym = sym;
    const int vec = 16 / (int)sizeof(T);
    o.vec_ok = (((uintptr_t)p & 15) == 0) && (ld % vec == 0);
    return o;
}

}  // namespace
