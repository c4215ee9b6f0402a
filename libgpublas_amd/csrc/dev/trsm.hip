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
    int rec(int m, int n, T alpha, const T *a, T *b)
    {
        int err;
        const int t = left ? m : n;
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
