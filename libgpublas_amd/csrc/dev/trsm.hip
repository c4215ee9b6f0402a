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
//  * left side:  one wave per right-hand-side column, lane i owns row i,
//                substitution broadcasts x_k with a wave shuffle;
//  * right side: one lane per row of B, the 64 unknowns of the row in
//                registers, triangle entries broadcast from LDS.
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

template <typename T>
__global__ __launch_bounds__(256) void trsm_leaf_left(int nb, int n, T alpha, const T *A, long lda,
                                                      int transN, int effLower, int nounit, T *B,
                                                      long ldb)
{
    __shared__ T Ts[NB][NB + 1];
    for (int idx = threadIdx.x; idx < NB * NB; idx += 256) {
        const int i = idx % NB, k = idx / NB;
        const bool in = i < nb && k < nb && (effLower ? i >= k : i <= k);
        Ts[i][k] = in ? opT(A, lda, transN, i, k) : T(0);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int j = blockIdx.x * 4 + wave; j < n; j += gridDim.x * 4) {
        T *col = B + (long)j * ldb;
        T b = lane < nb ? alpha * col[lane] : T(0);
        if (effLower) {
            for (int k = 0; k < nb; ++k) {
                T xk = __shfl(b, k, 64);
                if (nounit) xk = xk / Ts[k][k];
                if (lane == k) b = xk;
                else if (lane > k) b -= Ts[lane][k] * xk;
            }
        } else {
            for (int k = nb - 1; k >= 0; --k) {
                T xk = __shfl(b, k, 64);
                if (nounit) xk = xk / Ts[k][k];
                if (lane == k) b = xk;
                else if (lane < k) b -= Ts[lane][k] * xk;
            }
        }
        if (lane < nb) col[lane] = b;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void trsm_leaf_right(int m, int nb, T alpha, const T *A,
                                                       long lda, int transN, int effLower,
                                                       int nounit, T *B, long ldb)
{
    __shared__ T Ts[NB][NB + 1];
    for (int idx = threadIdx.x; idx < NB * NB; idx += 256) {
        const int k = idx % NB, j = idx / NB;
        const bool in = k < nb && j < nb && (effLower ? k >= j : k <= j);
        Ts[k][j] = in ? opT(A, lda, transN, k, j) : T(0);
    }
    __syncthreads();
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    T x[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) x[j] = j < nb ? alpha * B[i + (long)j * ldb] : T(0);
    if (effLower) {
        // X * L = B: x_j = (b_j - sum_{k>j} x_k L(k,j)) / L(j,j), j descending
#pragma unroll
        for (int j = NB - 1; j >= 0; --j) {
            if (j < nb) {
                T v = x[j];
#pragma unroll
                for (int k = j + 1; k < NB; ++k) v -= x[k] * Ts[k][j];
                if (nounit) v = v / Ts[j][j];
                x[j] = v;
            }
        }
    } else {
        // X * U = B: x_j = (b_j - sum_{k<j} x_k U(k,j)) / U(j,j), j ascending
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            if (j < nb) {
                T v = x[j];
#pragma unroll
                for (int k = 0; k < j; ++k) v -= x[k] * Ts[k][j];
                if (nounit) v = v / Ts[j][j];
                x[j] = v;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < NB; ++j)
        if (j < nb) B[i + (long)j * ldb] = x[j];
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
        if (left) {
            int blocks = (n + 3) / 4;
            if (blocks > 4096) blocks = 4096;
            PROF_BEGIN(s, sizeof(T) == 8 ? "dtrsm_leaf_left" : "strsm_leaf_left");
            hipLaunchKernelGGL(trsm_leaf_left<T>, dim3(blocks), dim3(256), 0, s, m, n, alpha, a,
                               (long)lda, transN, effLower, nounit, b, (long)ldb);
            PROF_END(s);
        } else {
            int blocks = (m + 255) / 256;
            PROF_BEGIN(s, sizeof(T) == 8 ? "dtrsm_leaf_right" : "strsm_leaf_right");
            hipLaunchKernelGGL(trsm_leaf_right<T>, dim3(blocks), dim3(256), 0, s, m, n, alpha, a,
                               (long)lda, transN, effLower, nounit, b, (long)ldb);
            PROF_END(s);
        }
        return (int)hipGetLastError();
    }

    // Solve on the sub-problem whose triangle starts at a (order `t`), with
    // right-hand sides at b (m x n, m = t for left, n = t for right).
    int rec(int m, int n, T alpha, const T *a, T *b)
    {
        int err;
        const int t = left ? m : n;
        if (t <= NB) return leaf(m, n, alpha, a, b);
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
