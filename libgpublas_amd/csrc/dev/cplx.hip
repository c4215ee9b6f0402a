// cplx.hip -- complex GEMM (ZGEMM / CGEMM) on gfx950 through the real MFMA
// GEMM (replaces cublasZgemm / cublasCgemm behind the reference's F77_gemm(z)
// and F77_gemm(c), blas_level3/gemm.cc:181-217; netlib semantics, 'C' is
// the conjugate transpose).
//
// C := alpha op(A) op(B) + beta C with interleaved (re, im) operands runs as
// ONE real product of m x 2n x 2k:
//   [Pr | Pi] = [Ar  Ai] . [[ Br  Bi ]
//                           [-Bi  Br ]]
// where Ar + i Ai = op(A) (m x k) and Br + i Bi = op(B) (k x n), conjugated
// for 'C'; so Pr = Ar Br - Ai Bi and Pi = Ar Bi + Ai Br, the 8mnk real flops
// of the complex product in one launch of the fast MFMA tiles (twice the
// columns and twice the depth of a real product of the same shape, which
// also fills the chip sooner).  Three passes around it, each one sweep of
// HBM: pack_a splits op(A)'s stored matrix into its two parts (kept in the
// stored orientation, so a transposed A stays a transposed real operand),
// pack_b builds the 2k x 2n block operand, and combine forms
// C = alpha (Pr + i Pi) + beta C (C is not read when beta = 0, as netlib).
// Each real entry of P is one k-ordered sum of 2k products (the real parts
// first, then the imaginary ones), a fixed order: results repeat bit for bit.
#include "devcommon.hpp"
#include "gpublas_dev.h"

int gpublas_gemm_d(hipStream_t s, char ta, char tb, int m, int n, int k, double alpha,
                   const double *A, int lda, const double *B, int ldb, double beta, double *C,
                   int ldc);
int gpublas_gemm_s(hipStream_t s, char ta, char tb, int m, int n, int k, float alpha,
                   const float *A, int lda, const float *B, int ldb, float beta, float *C,
                   int ldc);

namespace {

inline bool is_n(char c) { return c == 'N' || c == 'n'; }
inline bool is_c(char c) { return c == 'C' || c == 'c'; }

inline unsigned grid_for(long elems)
{
    const long b = (elems + 255) / 256;
    return (unsigned)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

// A's stored matrix (R x Cc, leading dimension lda in complex elements) into
// its real and imaginary parts: untransposed A as [Ar Ai] (R x 2Cc, ld R),
// a transposed one as [Ar; Ai] (2R x Cc, ld 2R); sgn = -1 conjugates.
template <typename T>
__global__ __launch_bounds__(256) void pack_a(int R, int Cc, const T *A, long lda, int trans,
                                              T sgn, T *out)
{
    const long total = (long)R * Cc;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const long r = e % R, c = e / R;
        const T re = A[2 * (r + c * lda)], im = sgn * A[2 * (r + c * lda) + 1];
        if (!trans) {
            out[r + c * R] = re;
            out[r + (Cc + c) * (long)R] = im;
        } else {
            out[r + c * 2L * R] = re;
            out[R + r + c * 2L * R] = im;
        }
    }
}

// The block operand Bt with op(Bt) = [[Br, Bi], [-Bi, Br]] (2k x 2n) from B's
// stored matrix (R x Cc, ld ldb complex).  Untransposed B (R = k, Cc = n):
// Bt itself, 2k x 2n, ld 2k.  Transposed B (R = n, Cc = k): Bt is stored
// 2n x 2k (ld 2n) as [[Br^T, -Bi^T], [Bi^T, Br^T]] of the stored parts.
template <typename T>
__global__ __launch_bounds__(256) void pack_b(int R, int Cc, const T *B, long ldb, int trans,
                                              T sgn, T *out)
{
    const long total = (long)R * Cc, ld = 2L * R;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const long r = e % R, c = e / R;
        const T re = B[2 * (r + c * ldb)], im = sgn * B[2 * (r + c * ldb) + 1];
        out[r + c * ld] = re;                 // Br (or Br^T)
        out[R + r + (Cc + c) * ld] = re;      // Br again, lower right
        if (!trans) {
            out[r + (Cc + c) * ld] = im;      // Bi, upper right
            out[R + r + c * ld] = -im;        // -Bi, lower left
        } else {
            out[R + r + c * ld] = im;         // Bi^T, lower left
            out[r + (Cc + c) * ld] = -im;     // -Bi^T, upper right
        }
    }
}

// C = alpha (Pr + i Pi) + beta C; P is m x 2n (ld m)
template <typename T>
__global__ __launch_bounds__(256) void combine(int m, int n, const T *P, T ar, T ai, T br, T bi,
                                               T *C, long ldc)
{
    const long total = (long)m * n;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const long i = e % m, j = e / m;
        const T pr = P[i + j * m], pi = P[i + (n + j) * (long)m];
        T re = ar * pr - ai * pi, im = ar * pi + ai * pr;
        T *c = C + 2 * (i + j * ldc);
        if (br != T(0) || bi != T(0)) {
            const T cr = c[0], ci = c[1];
            re += br * cr - bi * ci;
            im += br * ci + bi * cr;
        }
        c[0] = re;
        c[1] = im;
    }
}

// C := beta C (alpha = 0 or k = 0); exact zeros for beta = 0
template <typename T>
__global__ __launch_bounds__(256) void cscale(int m, int n, T br, T bi, T *C, long ldc)
{
    const long total = (long)m * n;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const long i = e % m, j = e / m;
        T *c = C + 2 * (i + j * ldc);
        if (br == T(0) && bi == T(0)) {
            c[0] = c[1] = T(0);
        } else {
            const T cr = c[0], ci = c[1];
            c[0] = br * cr - bi * ci;
            c[1] = br * ci + bi * cr;
        }
    }
}

int real_gemm(hipStream_t s, char ta, char tb, int m, int n, int k, const double *A, int lda,
              const double *B, int ldb, double *C, int ldc)
{
    return gpublas_gemm_d(s, ta, tb, m, n, k, 1.0, A, lda, B, ldb, 0.0, C, ldc);
}
int real_gemm(hipStream_t s, char ta, char tb, int m, int n, int k, const float *A, int lda,
              const float *B, int ldb, float *C, int ldc)
{
    return gpublas_gemm_s(s, ta, tb, m, n, k, 1.0f, A, lda, B, ldb, 0.0f, C, ldc);
}

template <typename T>
int cgemm_entry(void *stream, char transa, char transb, int m, int n, int k, T ar, T ai,
                const T *A, int lda, const T *B, int ldb, T br, T bi, T *C, int ldc)
{
    hipStream_t s = pick_stream(stream);
    if (m <= 0 || n <= 0) return 0;
    if ((ar == T(0) && ai == T(0)) || k <= 0) {
        if (br == T(1) && bi == T(0)) return 0;
        hipLaunchKernelGGL(cscale<T>, dim3(grid_for((long)m * n)), dim3(256), 0, s, m, n, br, bi,
                           C, (long)ldc);
        return (int)hipGetLastError();
    }
    // the real product is m x 2n x 2k; its int dimensions must fit
    if (n > 0x3fffffff || k > 0x3fffffff) return (int)hipErrorInvalidValue;
    const bool ntA = is_n(transa), ntB = is_n(transb);
    const int Ra = ntA ? m : k, Ca = ntA ? k : m;  // A's stored shape
    const int Rb = ntB ? k : n, Cb = ntB ? n : k;  // B's stored shape
    const size_t na = 2 * (size_t)m * k, nb = 4 * (size_t)k * n, np = 2 * (size_t)m * n;
    T *W = (T *)gpublas_dev_alloc((na + nb + np) * sizeof(T));
    if (!W) return (int)hipErrorOutOfMemory;
    T *At = W, *Bt = W + na, *P = W + na + nb;
    PROF_BEGIN(s, sizeof(T) == 8 ? "zgemm_pack" : "cgemm_pack");
    hipLaunchKernelGGL(pack_a<T>, dim3(grid_for((long)Ra * Ca)), dim3(256), 0, s, Ra, Ca, A,
                       (long)lda, ntA ? 0 : 1, is_c(transa) ? T(-1) : T(1), At);
    hipLaunchKernelGGL(pack_b<T>, dim3(grid_for((long)Rb * Cb)), dim3(256), 0, s, Rb, Cb, B,
                       (long)ldb, ntB ? 0 : 1, is_c(transb) ? T(-1) : T(1), Bt);
    PROF_END(s);
    int err = (int)hipGetLastError();
    if (!err)
        err = real_gemm(s, ntA ? 'N' : 'T', ntB ? 'N' : 'T', m, 2 * n, 2 * k, At,
                        ntA ? m : 2 * k, Bt, ntB ? 2 * k : 2 * n, P, m);
    if (!err) {
        PROF_BEGIN(s, sizeof(T) == 8 ? "zgemm_combine" : "cgemm_combine");
        hipLaunchKernelGGL(combine<T>, dim3(grid_for((long)m * n)), dim3(256), 0, s, m, n,
                           (const T *)P, ar, ai, br, bi, C, (long)ldc);
        PROF_END(s);
        err = (int)hipGetLastError();
    }
    dev_release_after(W, s);
    return err;
}

}  // namespace

extern "C" {

int gpublas_dev_zgemm(void *stream, char transa, char transb, int m, int n, int k,
                      double alpha_re, double alpha_im, const double *d_A, int lda,
                      const double *d_B, int ldb, double beta_re, double beta_im, double *d_C,
                      int ldc)
{
    return cgemm_entry<double>(stream, transa, transb, m, n, k, alpha_re, alpha_im, d_A, lda, d_B,
                               ldb, beta_re, beta_im, d_C, ldc);
}

int gpublas_dev_cgemm(void *stream, char transa, char transb, int m, int n, int k,
                      float alpha_re, float alpha_im, const float *d_A, int lda, const float *d_B,
                      int ldb, float beta_re, float beta_im, float *d_C, int ldc)
{
    return cgemm_entry<float>(stream, transa, transb, m, n, k, alpha_re, alpha_im, d_A, lda, d_B,
                              ldb, beta_re, beta_im, d_C, ldc);
}

}  // extern "C"
