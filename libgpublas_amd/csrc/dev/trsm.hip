// trsm.hip -- triangular solve with many right-hand sides (DTRSM/STRSM) on
// gfx950, all 16 side/uplo/trans/diag forms.
//
// Replaces the vendor call at blas_level3/trsm.cc:52-72 (cublasDtrsm).
//
// Recursive blocking: the triangle is split in two halves (rounded to the
// 64-row leaf size); one half is solved recursively, the off-diagonal block
// is applied to the other half's right-hand sides with the MFMA GEMM of
// gemm.hip (alpha = -1, beta = the solve's alpha), then the second half is
// solved.  So all but O(m * 64 * n) of the m^2 n flops run in large GEMMs.
// The 64 x 64 diagonal blocks are solved in FP64 (FP32 for STRSM) by leaf
// kernels that keep the triangle in LDS:
//  * one leaf kernel for both sides (the right side solved as the left
//    side of the transposed system): lane i owns unknown i of two
//    right-hand sides per wave, x_k broadcast from lane k with v_readlane.
#include "devcommon.hpp"
#include "gpublas_dev.h"

int gpublas_gemm_d(hipStream_t s, char ta, char tb, int m, int n, int k, double alpha,
                   const double *A, int lda, const double *B, int ldb, double beta, double *C,
                   int ldc);
int gpublas_gemm_s(hipStream_t s, char ta, char tb, int m, int n, int k, float alpha,
                   const float *A, int lda, const float *B, int ldb, float beta, float *C, int ldc);

namespace {

constexpr int NB = 64;

inline int gemm_t(hipStream_t s, char ta, char tb, int m, int n, int k, double alpha,
                  const double *A, int lda, const double *B, int ldb, double beta, double *C,
                  int ldc)
{
    return gpublas_gemm_d(s, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
inline int gemm_t(hipStream_t s, char ta, char tb, int m, int n, int k, float alpha,
                  const float *A, int lda, const float *B, int ldb, float beta, float *C, int ldc)
{
    return gpublas_gemm_s(s, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

// op(T)(i, k) of the nb x nb diagonal block starting at A.
template <typename T>
__device__ __forceinline__ T opT(const T *A, long lda, int transN, int i, int k)
{
    return transN ? A[i + (long)k * lda] : A[k + (long)i * lda];
}

// x from lane l of the wave (l uniform): two v_readlane for a double
__device__ __forceinline__ double lane_value(double v, int l)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ float lane_value(float v, int l)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}

// Solve of a 64 x 64 triangle, op(T) X = alpha B, for n right-hand sides:
// element s of right-hand side r is B[s*sstride + r*rstride] (left side:
// sstride 1, rstride ldb; right side, solved as op(T)^T X^T = alpha B^T:
// sstride ldb, rstride 1, with transN and the triangle flipped).  Lane i owns
// unknown i of LPW right-hand sides per wave (four waves per workgroup share
// the triangle in LDS).  Step k takes b_k of each right-hand side from lane k
// with v_readlane, forms x_k = b_k / T_kk as b_k * (1/T_kk) plus one Newton
// correction (three dependent operations instead of a full division on the
// step-to-step chain), and every later unknown subtracts x_k * T(i, k).  The
// LPW right-hand sides are independent chains, interleaved.  On the left side
// the update order per column is the reference's (netlib DTRSM, left side:
// B(k) /= A(k,k); B(i) -= B(k)*A(i,k)); on the right side the reference
// accumulates each unknown as a dot product instead, so sums are ordered
// differently.
constexpr int LPW = 2;
template <typename T, bool LOWER, bool NOUNIT>
__global__ __launch_bounds__(256) void trsm_leaf(int nb, int n, T alpha, const T *__restrict__ A,
                                                 long lda, int transN, T *__restrict__ B,
                                                 long sstride, long rstride)
{
    __shared__ T Ts[NB][NB + 1];  // Ts[i][k] = op(T)(i, k) (identity outside nb)
    __shared__ T rd[NB];          // 1 / op(T)(k, k)
    // op(T)(i, k) = A[i + k*lda] (transN) or A[k + i*lda]: consecutive
    // threads run along the contiguous index.  All 16 loads of a thread are
    // issued before the first LDS store, so their latency is paid once.
    constexpr int PER = NB * NB / 256;
    T tv[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int idx = threadIdx.x + 256 * u;
        const int i = transN ? idx % NB : idx / NB, k = transN ? idx / NB : idx % NB;
        const bool in = i < nb && k < nb && (LOWER ? i >= k : i <= k);
        tv[u] = in ? opT(A, lda, transN, i, k) : (i == k ? T(1) : T(0));
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int idx = threadIdx.x + 256 * u;
        const int i = transN ? idx % NB : idx / NB, k = transN ? idx / NB : idx % NB;
        Ts[i][k] = tv[u];
    }
    __syncthreads();
    if (threadIdx.x < NB) rd[threadIdx.x] = T(1) / Ts[threadIdx.x][threadIdx.x];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int jb = (blockIdx.x * 4 + wave) * LPW;
    T b[LPW];
#pragma unroll
    for (int c = 0; c < LPW; ++c) {
        const int j = jb + c;
        b[c] = (j < n && lane < nb) ? alpha * B[lane * sstride + j * rstride] : T(0);
    }
    for (int st = 0; st < NB; ++st) {
        const int k = LOWER ? st : NB - 1 - st;
        const T t = Ts[lane][k];
        const T r = rd[k], d = Ts[k][k];
        const bool below = LOWER ? lane > k : lane < k;
#pragma unroll
        for (int c = 0; c < LPW; ++c) {
            const T bk = lane_value(b[c], k);
            T x = bk;
            if (NOUNIT) {
                const T x0 = bk * r;
                x = fma(fma(-x0, d, bk), r, x0);
            }
            b[c] = lane == k ? x : below ? fma(-x, t, b[c]) : b[c];
        }
    }
#pragma unroll
    for (int c = 0; c < LPW; ++c) {
        const int j = jb + c;
        if (j < n && lane < nb) B[lane * sstride + j * rstride] = b[c];
    }
}

// Solve of an nb x nb triangle (nb <= 64 * R) for n right-hand sides in one
// launch: the base of the recursion, which otherwise solves four 64 x 64
// leaves and three small GEMM updates per 256 rows, each a launch of its own
// with too few tiles to fill the chip.  Lane i of a wave owns rows i, i + 64,
// ... (R groups) of LPW right-hand sides.  The triangle goes through LDS one
// 64-column panel at a time (the panel's rows from its diagonal block down
// for a lower triangle, up for an upper one); within a panel, step k takes x_k
// from lane k of the panel's row group (v_readlane, reciprocal plus one Newton
// correction as in trsm_leaf) and updates the panel's later rows and every
// row group beyond the panel.  Same element order per right-hand side as the
// leaves and GEMM updates of the recursion it replaces, up to the order of
// the sums over k within a 64-column panel.
constexpr int BIG_R = 4;                  // row groups: orders up to 256
// Right-hand sides per wave in the base kernel.  Every workgroup streams the
// whole triangle through LDS (133 KB: one workgroup per CU), so fewer, wider
// workgroups pay for it fewer times: at 8192 x 4096 the base launches took
// 98 us (2 per wave, 512 workgroups), 68 us (4, 256) and 104 us (8, 128).
constexpr int TRSM_LPWB = 4;
template <typename T, bool LOWER, bool NOUNIT, int LPWB>
__global__ __launch_bounds__(256) void trsm_leaf_big(int nb, int n, T alpha,
                                                     const T *__restrict__ A, long lda,
                                                     int transN, T *__restrict__ B, long sstride,
                                                     long rstride)
{
    __shared__ T Ts[NB * BIG_R][NB + 1];  // the current panel: Ts[row][k]
    __shared__ T rd[NB];                  // 1 / op(T)(kg, kg) of the panel
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int jb = (blockIdx.x * 4 + wave) * LPWB;
    T b[LPWB][BIG_R];
#pragma unroll
    for (int c = 0; c < LPWB; ++c)
#pragma unroll
        for (int g = 0; g < BIG_R; ++g) {
            const int j = jb + c, r = NB * g + lane;
            b[c][g] = (j < n && r < nb) ? alpha * B[r * sstride + j * rstride] : T(0);
        }
#pragma unroll
    for (int pp = 0; pp < BIG_R; ++pp) {
        const int p = LOWER ? pp : BIG_R - 1 - pp;
        if (NB * p >= nb) continue;  // uniform: panels past the order are empty
        // rows the panel feeds: lower from NB*p down, upper from NB*p+NB-1 up
        const int r_lo = LOWER ? NB * p : 0, r_hi = LOWER ? NB * BIG_R : NB * p + NB;
        __syncthreads();  // the previous panel's readers are done
        for (int e0 = r_lo * NB; e0 < r_hi * NB; e0 += 256 * 16) {
            T tv[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int e = e0 + threadIdx.x + 256 * u;
                // consecutive threads along the contiguous index of op(T)
                const int r = transN ? r_lo + (e - r_lo * NB) % (r_hi - r_lo)
                                     : r_lo + (e - r_lo * NB) / NB;
                const int kk = transN ? (e - r_lo * NB) / (r_hi - r_lo) : (e - r_lo * NB) % NB;
                const int kg = NB * p + kk;
                const bool in = e < r_hi * NB && r < nb && kg < nb && (LOWER ? r >= kg : r <= kg);
                const long off = transN ? r + (long)kg * lda : kg + (long)r * lda;
                const T v = A[in ? off : 0];  // always a valid address: no branch per load
                tv[u] = in ? v : (r == kg ? T(1) : T(0));
            }
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int e = e0 + threadIdx.x + 256 * u;
                if (e < r_hi * NB) {
                    const int r = transN ? r_lo + (e - r_lo * NB) % (r_hi - r_lo)
                                         : r_lo + (e - r_lo * NB) / NB;
                    const int kk =
                        transN ? (e - r_lo * NB) / (r_hi - r_lo) : (e - r_lo * NB) % NB;
                    Ts[r][kk] = tv[u];
                }
            }
        }
        __syncthreads();
        if (threadIdx.x < NB) rd[threadIdx.x] = T(1) / Ts[NB * p + threadIdx.x][threadIdx.x];
        __syncthreads();
        for (int st = 0; st < NB; ++st) {
            const int k = LOWER ? st : NB - 1 - st;
            const T t = Ts[NB * p + lane][k];
            const T r = rd[k], d = Ts[NB * p + k][k];
            const bool below = LOWER ? lane > k : lane < k;
            T tg[BIG_R];
#pragma unroll
            for (int g = 0; g < BIG_R; ++g)
                if (LOWER ? g > p : g < p) tg[g] = Ts[NB * g + lane][k];
#pragma unroll
            for (int c = 0; c < LPWB; ++c) {
                const T bk = lane_value(b[c][p], k);
                T x = bk;
                if (NOUNIT) {
                    const T x0 = bk * r;
                    x = fma(fma(-x0, d, bk), r, x0);
                }
                b[c][p] = lane == k ? x : below ? fma(-x, t, b[c][p]) : b[c][p];
#pragma unroll
                for (int g = 0; g < BIG_R; ++g)
                    if (LOWER ? g > p : g < p) b[c][g] = fma(-x, tg[g], b[c][g]);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < LPWB; ++c)
#pragma unroll
        for (int g = 0; g < BIG_R; ++g) {
            const int j = jb + c, r = NB * g + lane;
            if (j < n && r < nb) B[r * sstride + j * rstride] = b[c][g];
        }
}

// FP64 base solve on the matrix cores: op(T) X = alpha B for an order of at
// most 256, one 512-thread workgroup per 16 right-hand sides, the 16-column
// panel of B held in LDS, 16-row blocks.
//  * Prologue: the inverses of the 16 diagonal blocks are formed in LDS,
//    four blocks per wave, one lane per column (a substitution with the unit
//    vector; reciprocal of each diagonal entry, computed off the chain),
//    while the panel's loads are still in flight.
//  * Wave w keeps its blocks w and w + 8 in registers in accumulator layout.
//    Step s: X_I (I the s-th block in solve order) is in the panel; every
//    wave subtracts op(T)[J, I] X_I from its blocks J beyond I (two
//    independent chains of two v_mfma_f64_16x16x4), the owner of the next
//    block first, and that owner then forms X = inv(T_JJ) B_J with four
//    independent MFMAs that take the update's accumulator registers directly
//    as their B operand (the accumulator layout of one is the operand layout
//    of the other) and stores X into the panel; one LDS-only barrier per step.
//    The op(T) fragments are fetched two steps ahead (unrolled case).
// The sequential part per block is a few MFMAs and a barrier instead of a
// 16-step substitution (multiplying by the inverse of a 16 x 16 diagonal
// block is the usual GPU TRSM; exact for integer data, and within the
// usual triangular-solve error bound for well-conditioned diagonal blocks).
// Measured at order 256 with 4096 right-hand sides: 26 us per launch against
// 68 us for the substitution kernel (trsm_leaf_big, still the FP32 base).
constexpr int MB_MAX = 256;   // orders up to 256: 16 blocks
constexpr int MB_WAVES = 8;   // two blocks per wave
// FULL = 16: the order is 241..256 and the steps are unrolled (with no loop
// back edge the compiler waits for each prefetch only where it is used);
// FULL = 0: a loop over the steps for other orders.
template <bool LOWER, bool NOUNIT, int FULL>
__global__ __launch_bounds__(512) void trsm_base_mfma(int nbr, int n, double alpha,
                                                      const double *__restrict__ A, long lda,
                                                      int transN, double *__restrict__ B,
                                                      long sstride, long rstride)
{
    __shared__ double Bs[MB_MAX][17];      // Bs[row][rhs]
    __shared__ double Ls[16][16][17];      // diagonal blocks, then their inverses
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int q0 = lane >> 4, cl = lane & 15;
    const int c0 = blockIdx.x * 16;
    const int nb = FULL ? FULL : (nbr + 15) / 16;
    const int rows = nb * 16;
    // Prologue.  The diagonal blocks' loads are issued first and the panel's
    // after them, so the inversion waits only for the former (loads complete
    // in order) and runs while the panel is still in flight.
    constexpr int PER = MB_MAX * 16 / 512;
    double v[PER], d[PER];
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int idx = threadIdx.x + 512 * e;
        const int bk = idx >> 8, i = (idx >> 4) & 15, k = idx & 15;
        const int gi = 16 * bk + i, gk = 16 * bk + k;
        const bool tri = bk < nb && gi < nbr && gk < nbr && (LOWER ? k < i : k > i);
        d[e] = A[tri ? (transN ? gi + (long)gk * lda : gk + (long)gi * lda)
                     : (NOUNIT && bk < nb && gi < nbr && i == k ? gi + (long)gi * lda : 0)];
    }
    // alpha B: contiguous index fastest (rows on the left side, right-hand
    // sides on the right)
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int idx = threadIdx.x + 512 * e;
        const int r = sstride == 1 ? idx % rows : idx / 16;
        const int c = sstride == 1 ? idx / rows : idx % 16;
        const bool in = idx < rows * 16 && r < nbr && c0 + c < n;
        v[e] = B[in ? r * sstride + (long)(c0 + c) * rstride : 0];
    }
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int idx = threadIdx.x + 512 * e;
        const int bk = idx >> 8, i = (idx >> 4) & 15, k = idx & 15;
        const int gi = 16 * bk + i, gk = 16 * bk + k;
        const bool tri = bk < nb && gi < nbr && gk < nbr && (LOWER ? k < i : k > i);
        // identity past the order, unit diagonal when asked
        Ls[bk][i][k] = tri ? d[e] : i == k ? ((NOUNIT && bk < nb && gi < nbr) ? d[e] : 1.0) : 0.0;
    }
    lds_barrier();  // the diagonal blocks are in LDS (the panel's loads still in flight)
    if (wave < 4) {
        // inverse of diagonal block 4 * wave + q0, column cl of it in this lane
        // (the reciprocals of the diagonal are off the substitution's chain)
        const int bk = 4 * wave + q0;
        if (bk < nb) {
            double y[16];
#pragma unroll
            for (int ii = 0; ii < 16; ++ii) {
                const int i = LOWER ? ii : 15 - ii;
                const double rc = 1.0 / Ls[bk][i][i];  // independent of the chain
                double sacc = i == cl ? 1.0 : 0.0;
#pragma unroll
                for (int kk = 0; kk < ii; ++kk) {
                    const int k = LOWER ? kk : 15 - kk;
                    sacc = fma(-Ls[bk][i][k], y[k], sacc);
                }
                y[i] = sacc * rc;
            }
            // every lane's reads of the block are done before it is overwritten
            __builtin_amdgcn_s_waitcnt(vm_wait(63) & ~(15 << 8));
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int i = 0; i < 16; ++i) Ls[bk][i][cl] = y[i];
        }
    }
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int idx = threadIdx.x + 512 * e;
        const int r = sstride == 1 ? idx % rows : idx / 16;
        const int c = sstride == 1 ? idx / rows : idx % 16;
        const bool in = idx < rows * 16 && r < nbr && c0 + c < n;
        if (idx < rows * 16) Bs[r][c] = in ? alpha * v[e] : 0.0;
    }
    // op(T)[J, I] in MFMA A-operand layout for the wave's blocks J = wave + 8 s:
    // frag[s][u] = op(T)(16 J + cl, 16 I + 4 u + q0) where it lies in the block's
    // part of the triangle (in_tri).  Indices are clamped into the stored
    // nbr x nbr square, so every load has a valid address and no condition (a
    // branch here makes the compiler wait for the loads where it joins); the
    // mask is applied where the value is used.
    auto in_tri = [&](int I, int sb, int u) {
        const int J = wave + MB_WAVES * sb;
        const int i = 16 * J + cl, k = 16 * I + 4 * u + q0;
        return J < nb && (LOWER ? J > I : J < I) && i < nbr && k < nbr;
    };
    auto fetch = [&](int I, double (&frag)[2][4]) {
#pragma unroll
        for (int sb = 0; sb < 2; ++sb)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int i = min(16 * (wave + MB_WAVES * sb) + cl, nbr - 1);
                const int k = min(16 * I + 4 * u + q0, nbr - 1);
                frag[sb][u] = A[transN ? i + (long)k * lda : k + (long)i * lda];
            }
    };
    // The wave's two blocks stay in registers in accumulator layout (bacc[s]:
    // rows 16 J + q0 + 4 r, column cl); only solved blocks go to the panel.
    v4d bacc[2];
    // X_J = inv(T_JJ) B_J into the panel
    auto solve = [&](int sb) {
        const int J = wave + MB_WAVES * sb;
        // four independent k-chunk products (no accumulator chain), then summed
        const v4d z = {0.0, 0.0, 0.0, 0.0};
        v4d p[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            p[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(Ls[J][cl][4 * u + q0], bacc[sb][u], z, 0,
                                                        0, 0);
        const v4d x = (p[0] + p[1]) + (p[2] + p[3]);
#pragma unroll
        for (int r = 0; r < 4; ++r) Bs[16 * J + q0 + 4 * r][cl] = x[r];
    };
    // step st: X_I is in the panel (I = the st-th block in solve order);
    // updates with it, Inext solved.  `ahead` (fragments of the step `dist`
    // steps later) is fetched now, a step's duration being shorter than the
    // loads' latency.
    auto step = [&](int st, double (&cur)[2][4], double (&ahead)[2][4], int dist) {
        const int I = LOWER ? st : nb - 1 - st;
        const int Inext = LOWER ? I + 1 : I - 1;
        if (st + dist < nb - 1) fetch(LOWER ? st + dist : nb - 1 - (st + dist), ahead);
        double xo[4];  // X_I in B-operand layout: row 16 I + 4 u + q0, column cl
#pragma unroll
        for (int u = 0; u < 4; ++u) xo[u] = Bs[16 * I + 4 * u + q0][cl];
        const int first = (wave + MB_WAVES == Inext) ? 1 : 0;  // the next block first
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int sb = q ? 1 - first : first;
            const int J = wave + MB_WAVES * sb;
            if (J < nb && (LOWER ? J > I : J < I)) {
                // B_J -= op(T)[J, I] X_I: two independent chains of two k chunks
                const v4d z = {0.0, 0.0, 0.0, 0.0};
                v4d p0 = __builtin_amdgcn_mfma_f64_16x16x4f64(
                    in_tri(I, sb, 0) ? -cur[sb][0] : 0.0, xo[0], bacc[sb], 0, 0, 0);
                v4d p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(
                    in_tri(I, sb, 2) ? -cur[sb][2] : 0.0, xo[2], z, 0, 0, 0);
                p0 = __builtin_amdgcn_mfma_f64_16x16x4f64(in_tri(I, sb, 1) ? -cur[sb][1] : 0.0,
                                                         xo[1], p0, 0, 0, 0);
                p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(in_tri(I, sb, 3) ? -cur[sb][3] : 0.0,
                                                         xo[3], p1, 0, 0, 0);
                bacc[sb] = p0 + p1;
                if (J == Inext) solve(sb);
            }
        }
        lds_barrier();  // X_Inext in the panel
    };
    // fragment buffers: a ring of three (two steps ahead) when unrolled, two otherwise
    double fr[3][2][4];
    fetch(LOWER ? 0 : nb - 1, fr[0]);
    if constexpr (FULL > 0) fetch(LOWER ? 1 : nb - 2, fr[1]);
    __syncthreads();  // the inverses are in LDS
#pragma unroll
    for (int sb = 0; sb < 2; ++sb) {
        const int J = min(wave + MB_WAVES * sb, 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) bacc[sb][r] = Bs[16 * J + q0 + 4 * r][cl];
    }
    {
        const int I = LOWER ? 0 : nb - 1;
        if (wave == I % MB_WAVES) solve(I / MB_WAVES);
        lds_barrier();
    }
    if constexpr (FULL > 0) {
#pragma unroll
        for (int st = 0; st < FULL - 1; ++st) step(st, fr[st % 3], fr[(st + 2) % 3], 2);
    } else {
        for (int st = 0; st < nb - 1; st += 2) {
            step(st, fr[0], fr[1], 1);
            if (st + 1 < nb - 1) step(st + 1, fr[1], fr[0], 1);
        }
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < rows * 16; idx += 512) {
        const int r = sstride == 1 ? idx % rows : idx / 16;
        const int c = sstride == 1 ? idx / rows : idx % 16;
        if (r < nbr && c0 + c < n) B[r * sstride + (long)(c0 + c) * rstride] = Bs[r][c];
    }
}

template <typename T>
__global__ void zero_kernel(int m, int n, T *B, long ldb)
{
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long total = (long)m * n;
    for (; idx < total; idx += (long)gridDim.x * blockDim.x) B[idx % m + (idx / m) * ldb] = T(0);
}

inline int split(int m)
{
    int h = (m / 2 + NB - 1) / NB * NB;
    if (h >= m) h = m - NB;
    return h;
}

template <typename T>
struct Solve {
    hipStream_t s;
    int left, effLower, transN, nounit;
    const T *A;
    int lda;
    T *B;
    int ldb;

    int leaf(int m, int n, T alpha, const T *a, T *b)
    {
        // right side: op(T)^T X^T = alpha B^T, the triangle of op(T)^T is
        // op(T) read the other way and of the other shape
        const bool lower = left ? effLower : !effLower;
        const int tn = left ? transN : !transN;
        const int nrhs = left ? n : m, order = left ? m : n;
        const long ss = left ? 1 : ldb, rs = left ? ldb : 1;
        const unsigned blocks = (unsigned)ceil_div(nrhs, 4 * LPW);
        auto k = lower ? (nounit ? trsm_leaf<T, true, true> : trsm_leaf<T, true, false>)
                       : (nounit ? trsm_leaf<T, false, true> : trsm_leaf<T, false, false>);
        PROF_BEGIN(s, sizeof(T) == 8 ? (left ? "dtrsm_leaf_left" : "dtrsm_leaf_right")
                                     : (left ? "strsm_leaf_left" : "strsm_leaf_right"));
        hipLaunchKernelGGL(k, dim3(blocks), dim3(256), 0, s, order, nrhs, alpha, a, (long)lda, tn, b,
                           ss, rs);
        PROF_END(s);
        return (int)hipGetLastError();
    }

    // orders NB+1 .. NB*BIG_R in one launch (trsm_leaf_big)
    template <int LPWB>
    int leaf_big(int m, int n, T alpha, const T *a, T *b)
    {
        const bool lower = left ? effLower : !effLower;
        const int tn = left ? transN : !transN;
        const int nrhs = left ? n : m, order = left ? m : n;
        const long ss = left ? 1 : ldb, rs = left ? ldb : 1;
        const unsigned blocks = (unsigned)ceil_div(nrhs, 4 * LPWB);
        auto k = lower ? (nounit ? trsm_leaf_big<T, true, true, LPWB>
                                 : trsm_leaf_big<T, true, false, LPWB>)
                       : (nounit ? trsm_leaf_big<T, false, true, LPWB>
                                 : trsm_leaf_big<T, false, false, LPWB>);
        PROF_BEGIN(s, sizeof(T) == 8 ? "dtrsm_base" : "strsm_base");
        hipLaunchKernelGGL(k, dim3(blocks), dim3(256), 0, s, order, nrhs, alpha, a, (long)lda, tn,
                           b, ss, rs);
        PROF_END(s);
        return (int)hipGetLastError();
    }

    // Solve on the sub-problem whose triangle starts at a (order `t`), with
    // right-hand sides at b (m x n, m = t for left, n = t for right).
    // FP64 orders up to MB_MAX in one launch on the matrix cores (trsm_base_mfma)
    int base_mfma(int m, int n, T alpha, const T *a, T *b)
    {
        const bool lower = left ? effLower : !effLower;
        const int tn = left ? transN : !transN;
        const int nrhs = left ? n : m, order = left ? m : n;
        const long ss = left ? 1 : ldb, rs = left ? ldb : 1;
        const unsigned blocks = (unsigned)ceil_div(nrhs, 16);
        const bool full = order > MB_MAX - 16;
        auto k = full ? (lower ? (nounit ? trsm_base_mfma<true, true, 16>
                                         : trsm_base_mfma<true, false, 16>)
                               : (nounit ? trsm_base_mfma<false, true, 16>
                                         : trsm_base_mfma<false, false, 16>))
                      : (lower ? (nounit ? trsm_base_mfma<true, true, 0>
                                         : trsm_base_mfma<true, false, 0>)
                               : (nounit ? trsm_base_mfma<false, true, 0>
                                         : trsm_base_mfma<false, false, 0>));
        PROF_BEGIN(s, "dtrsm_base");
        hipLaunchKernelGGL(k, dim3(blocks), dim3(512), 0, s, order, nrhs, (double)alpha,
                           (const double *)a, (long)lda, tn, (double *)b, ss, rs);
        PROF_END(s);
        return (int)hipGetLastError();
    }

    int rec(int m, int n, T alpha, const T *a, T *b)
    {
        int err;
        const int t = left ? m : n;
        if constexpr (sizeof(T) == 8)
            if (t <= MB_MAX) return base_mfma(m, n, alpha, a, b);
        if (t <= NB) return leaf(m, n, alpha, a, b);
        if (t <= NB * BIG_R) return leaf_big<TRSM_LPWB>(m, n, alpha, a, b);
        const int t1 = split(t), t2 = t - t1;
        const T *a11 = a, *a22 = a + t1 + (long)t1 * lda;
        const T *a21 = a + t1, *a12 = a + (long)t1 * lda;  // stored blocks
        const char N = 'N', Tt = 'T';
        if (left) {
            T *b1 = b, *b2 = b + t1;
            if (effLower) {
                // op(A) = [[T11,0],[T21,T22]], T21 = A21 (N) or A12^T (T)
                if ((err = rec(t1, n, alpha, a11, b1))) return err;
                if (transN) err = gemm_t(s, N, N, t2, n, t1, T(-1), a21, lda, b1, ldb, alpha, b2, ldb);
                else err = gemm_t(s, Tt, N, t2, n, t1, T(-1), a12, lda, b1, ldb, alpha, b2, ldb);
                if (err) return err;
                return rec(t2, n, T(1), a22, b2);
            }
            // op(A) = [[T11,T12],[0,T22]], T12 = A12 (N) or A21^T (T)
            if ((err = rec(t2, n, alpha, a22, b2))) return err;
            if (transN) err = gemm_t(s, N, N, t1, n, t2, T(-1), a12, lda, b2, ldb, alpha, b1, ldb);
            else err = gemm_t(s, Tt, N, t1, n, t2, T(-1), a21, lda, b2, ldb, alpha, b1, ldb);
            if (err) return err;
            return rec(t1, n, T(1), a11, b1);
        }
        T *b1 = b, *b2 = b + (long)t1 * ldb;
        if (effLower) {
            // X [[T11,0],[T21,T22]] = alpha [B1,B2]
            if ((err = rec(m, t2, alpha, a22, b2))) return err;
            if (transN) err = gemm_t(s, N, N, m, t1, t2, T(-1), b2, ldb, a21, lda, alpha, b1, ldb);
            else err = gemm_t(s, N, Tt, m, t1, t2, T(-1), b2, ldb, a12, lda, alpha, b1, ldb);
            if (err) return err;
            return rec(m, t1, T(1), a11, b1);
        }
        // X [[T11,T12],[0,T22]] = alpha [B1,B2]
        if ((err = rec(m, t1, alpha, a11, b1))) return err;
        if (transN) err = gemm_t(s, N, N, m, t2, t1, T(-1), b1, ldb, a12, lda, alpha, b2, ldb);
        else err = gemm_t(s, N, Tt, m, t2, t1, T(-1), b1, ldb, a21, lda, alpha, b2, ldb);
        if (err) return err;
        return rec(m, t2, T(1), a22, b2);
    }
};

inline bool up(char c) { return c == 'U' || c == 'u'; }
inline bool isN(char c) { return c == 'N' || c == 'n'; }
inline bool isL(char c) { return c == 'L' || c == 'l'; }

template <typename T>
int trsm_entry(void *stream, char side, char uplo, char transa, char diag, int m, int n, T alpha,
               const T *A, int lda, T *B, int ldb)
{
    hipStream_t s = pick_stream(stream);
    if (m <= 0 || n <= 0) return 0;
    if (alpha == T(0)) {
        long total = (long)m * n;
        int blocks = (int)((total + 255) / 256);
        if (blocks > 8192) blocks = 8192;
        hipLaunchKernelGGL(zero_kernel<T>, dim3(blocks), dim3(256), 0, s, m, n, B, (long)ldb);
        return (int)hipGetLastError();
    }
    Solve<T> sv;
    sv.s = s;
    sv.left = isL(side);
    sv.transN = isN(transa);
    sv.effLower = (!up(uplo)) == sv.transN;
    sv.nounit = isN(diag);
    sv.A = A;
    sv.lda = lda;
    sv.B = B;
    sv.ldb = ldb;
    return sv.rec(m, n, alpha, A, B);
}

}  // namespace

extern "C" {

int gpublas_dev_dtrsm(void *stream, char side, char uplo, char transa, char diag, int m, int n,
                      double alpha, const double *d_A, int lda, double *d_B, int ldb)
{
    return trsm_entry<double>(stream, side, uplo, transa, diag, m, n, alpha, d_A, lda, d_B, ldb);
}

int gpublas_dev_strsm(void *stream, char side, char uplo, char transa, char diag, int m, int n,
                      float alpha, const float *d_A, int lda, float *d_B, int ldb)
{
    return trsm_entry<float>(stream, side, uplo, transa, diag, m, n, alpha, d_A, lda, d_B, ldb);
}

}  // extern "C"
