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
int gpublas_gemm_tri_d(hipStream_t s, char uplo, char ta, char tb, int n, int k, double alpha,
                       const double *A, int lda, const double *B, int ldb, double beta, double *C,
                       int ldc);
int gpublas_gemm_tri_s(hipStream_t s, char uplo, char ta, char tb, int n, int k, float alpha,
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

// ---- complex SYRK / HERK / SYR2K / HER2K ---------------------------------
// (syrk.cc:182-215, herk.cc:131-162, syr2k.cc:182-215, her2k.cc:141-178 of
// the reference).  With X = op(A), Y = op(B) (n x k; op = A, A^T or, for the
// Hermitian forms, A^H) the triangle of the result is two real products on
// the triangular tile grid, P_r = U V^T and P_i = U W^T, over stacked real
// operands of n x 2k (one product) or n x 4k (two products):
//   SYRK   X X^T               U = [Xr Xi]        V = [Xr -Xi]       W = [Xi Xr]
//   HERK   X X^H               U = [Xr Xi]        V = [Xr Xi]        W = [-Xi Xr]
//   SYR2K  X Y^T + Y X^T       U = [Xr Xi Yr Yi]  V = [Yr -Yi Xr -Xi] W = [Yi Yr Xi Xr]
//   HER2K  aX Y^H + conj(a) Y X^H, with X' = aX:
//                              U = [X'r X'i Yr Yi] V = [Yr Yi X'r X'i] W = [-Yi Yr -X'i X'r]
// and a last pass forms C = alpha P + beta C on the triangle (HER2K: P + beta
// C, alpha is in X'), the diagonal of a Hermitian result real (netlib).
enum { K_SYRK = 0, K_HERK = 1, K_SYR2K = 2, K_HER2K = 3 };

struct TriPack {  // column blocks of one stacked operand: sgn * part(src)
    int nb;
    int src[4];   // 0: x, 1: y
    int part[4];  // 0: real, 1: imaginary
    int sgn[4];
};

// element (i, p) of op(X), op = X (tx 0), X^T (1), X^H (2)
template <typename T>
__device__ __forceinline__ void op_elem(const T *X, long ldx, int tx, long i, long p, T &re, T &im)
{
    const long e = tx == 0 ? i + p * ldx : p + i * ldx;
    re = X[2 * e];
    im = tx == 2 ? -X[2 * e + 1] : X[2 * e + 1];
}

template <typename T>
__device__ __forceinline__ void put_blocks(const TriPack &q, T xr, T xi, T yr, T yi, long i, long p,
                                           int k, T *out, long ld)
{
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        if (b >= q.nb) break;
        const T v = q.src[b] ? (q.part[b] ? yi : yr) : (q.part[b] ? xi : xr);
        out[i + ((long)b * k + p) * ld] = q.sgn[b] < 0 ? -v : v;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void pack_tri(int n, int k, const T *X, long ldx, const T *Y,
                                                long ldy, int tx, T cr, T ci, TriPack u,
                                                TriPack v, TriPack w, T *U, T *V, T *W, long ld)
{
    const long total = (long)n * k;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const long i = e % n, p = e / n;
        T xr, xi, yr = T(0), yi = T(0);
        op_elem(X, ldx, tx, i, p, xr, xi);
        if (cr != T(1) || ci != T(0)) {  // X' = c X (HER2K's alpha)
            const T r = cr * xr - ci * xi;
            xi = cr * xi + ci * xr;
            xr = r;
        }
        if (Y) op_elem(Y, ldy, tx, i, p, yr, yi);
        put_blocks(u, xr, xi, yr, yi, i, p, k, U, ld);
        put_blocks(v, xr, xi, yr, yi, i, p, k, V, ld);
        put_blocks(w, xr, xi, yr, yi, i, p, k, W, ld);
    }
}

__device__ __forceinline__ bool in_tri(bool upper, long i, long j) { return upper ? i <= j : i >= j; }

// C = alpha (Pr + i Pi) + beta C on the triangle; herm: the diagonal's
// imaginary part is zero and C's own diagonal imaginary part is not read.
// scale_only: C = beta C (P not read).  beta = 0: C not read.
template <typename T>
__global__ __launch_bounds__(256) void combine_tri(int n, int upper, const T *Pr, const T *Pi,
                                                   long ldp, T ar, T ai, T br, T bi, int herm,
                                                   int scale_only, T *C, long ldc)
{
    const long total = (long)n * n;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const long i = e % n, j = e / n;
        if (!in_tri(upper != 0, i, j)) continue;
        T re = T(0), im = T(0);
        if (!scale_only) {
            const T pr = Pr[i + j * ldp], pi = Pi[i + j * ldp];
            re = ar * pr - ai * pi;
            im = ar * pi + ai * pr;
        }
        T *c = C + 2 * (i + j * ldc);
        if (br != T(0) || bi != T(0)) {
            const T cr = c[0], cim = (herm && i == j) ? T(0) : c[1];
            re += br * cr - bi * cim;
            im += br * cim + bi * cr;
        }
        c[0] = re;
        c[1] = (herm && i == j) ? T(0) : im;
    }
}

int tri_gemm(hipStream_t s, char uplo, int n, int k, const double *U, const double *V, long ld,
             double *P)
{
    return gpublas_gemm_tri_d(s, uplo, 'N', 'T', n, k, 1.0, U, (int)ld, V, (int)ld, 0.0, P, (int)ld);
}
int tri_gemm(hipStream_t s, char uplo, int n, int k, const float *U, const float *V, long ld,
             float *P)
{
    return gpublas_gemm_tri_s(s, uplo, 'N', 'T', n, k, 1.0f, U, (int)ld, V, (int)ld, 0.0f, P, (int)ld);
}

TriPack tp(int nb, const int (*spec)[3])
{
    TriPack q;
    q.nb = nb;
    for (int b = 0; b < 4; ++b) {
        q.src[b] = b < nb ? spec[b][0] : 0;
        q.part[b] = b < nb ? spec[b][1] : 0;
        q.sgn[b] = b < nb ? spec[b][2] : 1;
    }
    return q;
}

template <typename T>
int crk_entry(void *stream, int kind, char uplo, char trans, int n, int k, T ar, T ai,
              const T *A, int lda, const T *B, int ldb, T br, T bi, T *C, int ldc)
{
    hipStream_t s = pick_stream(stream);
    if (n <= 0) return 0;
    const bool herm = kind == K_HERK || kind == K_HER2K;
    const bool upper = uplo == 'U' || uplo == 'u';
    const bool scale_only = (ar == T(0) && ai == T(0)) || k <= 0;
    if (scale_only) {
        if (br == T(1) && bi == T(0)) return 0;
        hipLaunchKernelGGL(combine_tri<T>, dim3(grid_for((long)n * n)), dim3(256), 0, s, n,
                           upper ? 1 : 0, (const T *)nullptr, (const T *)nullptr, 0L, T(0), T(0),
                           br, bi, herm ? 1 : 0, 1, C, (long)ldc);
        return (int)hipGetLastError();
    }
    if (k > 0x1fffffff) return (int)hipErrorInvalidValue;
    const bool two = kind == K_SYR2K || kind == K_HER2K;
    const int tx = is_n(trans) ? 0 : (herm ? 2 : 1);
    const int kk = (two ? 4 : 2) * k;  // depth of the stacked real products
    const long ld = ((long)n + 7) / 8 * 8;
    const size_t opnd = (size_t)ld * kk, prod = (size_t)ld * n;
    T *Wk = (T *)gpublas_dev_alloc((3 * opnd + 2 * prod) * sizeof(T));
    if (!Wk) return (int)hipErrorOutOfMemory;
    T *U = Wk, *V = U + opnd, *W = V + opnd, *Pr = W + opnd, *Pi = Pr + prod;
    // {src, part, sign} per column block (x = op(A) or alpha op(A), y = op(B))
    static const int syrk_u[2][3] = {{0, 0, 1}, {0, 1, 1}}, syrk_v[2][3] = {{0, 0, 1}, {0, 1, -1}},
                     syrk_w[2][3] = {{0, 1, 1}, {0, 0, 1}};
    static const int herk_v[2][3] = {{0, 0, 1}, {0, 1, 1}}, herk_w[2][3] = {{0, 1, -1}, {0, 0, 1}};
    static const int r2_u[4][3] = {{0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}};
    static const int syr2_v[4][3] = {{1, 0, 1}, {1, 1, -1}, {0, 0, 1}, {0, 1, -1}};
    static const int syr2_w[4][3] = {{1, 1, 1}, {1, 0, 1}, {0, 1, 1}, {0, 0, 1}};
    static const int her2_v[4][3] = {{1, 0, 1}, {1, 1, 1}, {0, 0, 1}, {0, 1, 1}};
    static const int her2_w[4][3] = {{1, 1, -1}, {1, 0, 1}, {0, 1, -1}, {0, 0, 1}};
    TriPack pu, pv, pw;
    switch (kind) {
    case K_SYRK: pu = tp(2, syrk_u), pv = tp(2, syrk_v), pw = tp(2, syrk_w); break;
    case K_HERK: pu = tp(2, syrk_u), pv = tp(2, herk_v), pw = tp(2, herk_w); break;
    case K_SYR2K: pu = tp(4, r2_u), pv = tp(4, syr2_v), pw = tp(4, syr2_w); break;
    default: pu = tp(4, r2_u), pv = tp(4, her2_v), pw = tp(4, her2_w); break;
    }
    // HER2K folds alpha into X; the others apply it in the combine
    const T cr = kind == K_HER2K ? ar : T(1), ci = kind == K_HER2K ? ai : T(0);
    PROF_BEGIN(s, sizeof(T) == 8 ? "zrk_pack" : "crk_pack");
    hipLaunchKernelGGL(pack_tri<T>, dim3(grid_for((long)n * k)), dim3(256), 0, s, n, k, A,
                       (long)lda, two ? B : (const T *)nullptr, (long)ldb, tx, cr, ci, pu, pv, pw,
                       U, V, W, ld);
    PROF_END(s);
    int err = (int)hipGetLastError();
    if (!err) err = tri_gemm(s, uplo, n, kk, U, V, ld, Pr);
    if (!err) err = tri_gemm(s, uplo, n, kk, U, W, ld, Pi);
    if (!err) {
        const T er = kind == K_HER2K ? T(1) : ar, ei = kind == K_HER2K ? T(0) : ai;
        PROF_BEGIN(s, sizeof(T) == 8 ? "zrk_combine" : "crk_combine");
        hipLaunchKernelGGL(combine_tri<T>, dim3(grid_for((long)n * n)), dim3(256), 0, s, n,
                           upper ? 1 : 0, (const T *)Pr, (const T *)Pi, ld, er, ei, br, bi,
                           herm ? 1 : 0, 0, C, (long)ldc);
        PROF_END(s);
        err = (int)hipGetLastError();
    }
    dev_release_after(Wk, s);
    return err;
}

// ---- complex SYMM / HEMM (symm.cc:158-207, hemm.cc:120-153) ---------------
// The stored triangle of A (na x na) is expanded into a full complex matrix
// (the other triangle mirrored; conjugated, with a real diagonal, for HEMM),
// then the complex GEMM runs on it.
template <typename T>
__global__ __launch_bounds__(256) void expand_sym(int na, const T *A, long lda, int upper,
                                                  int herm, T *F)
{
    const long total = (long)na * na;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const long i = e % na, j = e / na;
        const bool stored = in_tri(upper != 0, i, j);
        const long src = stored ? i + j * lda : j + i * lda;
        T re = A[2 * src], im = A[2 * src + 1];
        if (herm) im = i == j ? T(0) : (stored ? im : -im);
        F[2 * e] = re;
        F[2 * e + 1] = im;
    }
}

template <typename T>
int csymm_entry(void *stream, int herm, char side, char uplo, int m, int n, T ar, T ai,
                const T *A, int lda, const T *B, int ldb, T br, T bi, T *C, int ldc)
{
    hipStream_t s = pick_stream(stream);
    if (m <= 0 || n <= 0) return 0;
    const bool left = side == 'L' || side == 'l';
    const int na = left ? m : n;
    if ((ar == T(0) && ai == T(0)))
        return cgemm_entry<T>(s, 'N', 'N', m, n, 0, ar, ai, A, lda, B, ldb, br, bi, C, ldc);
    T *F = (T *)gpublas_dev_alloc(2 * (size_t)na * na * sizeof(T));
    if (!F) return (int)hipErrorOutOfMemory;
    hipLaunchKernelGGL(expand_sym<T>, dim3(grid_for((long)na * na)), dim3(256), 0, s, na, A,
                       (long)lda, (uplo == 'U' || uplo == 'u') ? 1 : 0, herm, F);
    int err = (int)hipGetLastError();
    if (!err)
        err = left ? cgemm_entry<T>(s, 'N', 'N', m, n, m, ar, ai, F, m, B, ldb, br, bi, C, ldc)
                   : cgemm_entry<T>(s, 'N', 'N', m, n, n, ar, ai, B, ldb, F, n, br, bi, C, ldc);
    dev_release_after(F, s);
    return err;
}

// ---- complex TRMM (netlib ZTRMM; trmm.cc:177-215 is not built by the reference)
// The triangle is expanded (zeros outside it, ones on a unit diagonal), the
// complex GEMM forms alpha op(T) B (or alpha B op(T)) into a workspace, and
// the workspace is copied over B.
template <typename T>
__global__ __launch_bounds__(256) void expand_tri(int na, const T *A, long lda, int upper, int unit,
                                                  T *F)
{
    const long total = (long)na * na;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const long i = e % na, j = e / na;
        T re = T(0), im = T(0);
        if (i == j && unit) re = T(1);
        else if (in_tri(upper != 0, i, j)) re = A[2 * (i + j * lda)], im = A[2 * (i + j * lda) + 1];
        F[2 * e] = re;
        F[2 * e + 1] = im;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void copy_cols(int m, int n, const T *W, long ldw, T *B, long ldb)
{
    const long total = (long)m * n;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const long i = e % m, j = e / m;
        B[2 * (i + j * ldb)] = W[2 * (i + j * ldw)];
        B[2 * (i + j * ldb) + 1] = W[2 * (i + j * ldw) + 1];
    }
}

template <typename T>
int ctrmm_entry(void *stream, char side, char uplo, char transa, char diag, int m, int n, T ar,
                T ai, const T *A, int lda, T *B, int ldb)
{
    hipStream_t s = pick_stream(stream);
    if (m <= 0 || n <= 0) return 0;
    if (ar == T(0) && ai == T(0))
        return cgemm_entry<T>(s, 'N', 'N', m, n, 0, T(0), T(0), B, ldb, B, ldb, T(0), T(0), B, ldb);
    const bool left = side == 'L' || side == 'l';
    const int na = left ? m : n;
    T *F = (T *)gpublas_dev_alloc(2 * ((size_t)na * na + (size_t)m * n) * sizeof(T));
    if (!F) return (int)hipErrorOutOfMemory;
    T *Wm = F + 2 * (size_t)na * na;
    hipLaunchKernelGGL(expand_tri<T>, dim3(grid_for((long)na * na)), dim3(256), 0, s, na, A,
                       (long)lda, (uplo == 'U' || uplo == 'u') ? 1 : 0,
                       (diag == 'U' || diag == 'u') ? 1 : 0, F);
    int err = (int)hipGetLastError();
    if (!err)
        err = left ? cgemm_entry<T>(s, transa, 'N', m, n, m, ar, ai, F, m, B, ldb, T(0), T(0), Wm, m)
                   : cgemm_entry<T>(s, 'N', transa, m, n, n, ar, ai, B, ldb, F, n, T(0), T(0), Wm, m);
    if (!err) {
        hipLaunchKernelGGL(copy_cols<T>, dim3(grid_for((long)m * n)), dim3(256), 0, s, m, n,
                           (const T *)Wm, (long)m, B, (long)ldb);
        err = (int)hipGetLastError();
    }
    dev_release_after(F, s);
    return err;
}

// ---- complex TRSM (trsm.cc:175-215 of the reference; netlib ZTRSM) -------
// Recursive halving with the complex GEMM for the off-diagonal updates; a
// triangle of order <= CT_BASE is solved by substitution (x_t = (b_t - the
// solved terms, subtracted in solve order) / a_tt), one thread per right-hand
// side (left side: a column of B; right side: a row) holding its CT_BASE
// unknowns in registers, op(A) in LDS (padded with the identity up to
// CT_BASE, so a smaller triangle solves the same way with zeros after it).
// op(A)(r, c) reads A(r, c) ('N'), A(c, r) ('T') or conj(A(c, r)) ('C').
constexpr int CT_BASE = 32;
constexpr int CT_THREADS = 128;

template <typename T>
__device__ __forceinline__ void opa(const T *A, long lda, int tr, long r, long c, T &re, T &im)
{
    const long e = tr == 0 ? r + c * lda : c + r * lda;
    re = A[2 * e];
    im = tr == 2 ? -A[2 * e + 1] : A[2 * e + 1];
}

// op(A) (nb x nb, lower in solve order when FWD) X = alpha B (LEFT) or
// X op(A) = alpha B, for nrhs right-hand sides of B
template <typename T, bool LEFT, bool FWD>
__global__ __launch_bounds__(CT_THREADS) void ctrsm_base(int nb, int nrhs, int tr, int unit, T ar,
                                                         T ai, const T *A, long lda, T *B,
                                                         long ldb)
{
    __shared__ T Sr[CT_BASE][CT_BASE + 1], Si[CT_BASE][CT_BASE + 1];  // op(A)(r, c) at [r][c]
    for (int e = threadIdx.x; e < CT_BASE * CT_BASE; e += CT_THREADS) {
        const int r = e % CT_BASE, c = e / CT_BASE;
        T re = r == c ? T(1) : T(0), im = T(0);
        if (r < nb && c < nb && !(unit && r == c)) opa(A, lda, tr, r, c, re, im);
        Sr[r][c] = re;
        Si[r][c] = im;
    }
    __syncthreads();
    const long q = (long)blockIdx.x * CT_THREADS + threadIdx.x;  // the right-hand side
    if (q >= nrhs) return;
    // register t holds unknown row(t); the solve always walks t downwards, so
    // a forward solve keeps its unknowns in reverse order (one loop shape for
    // both directions)
    auto row = [](int t) { return FWD ? CT_BASE - 1 - t : t; };
    T xr[CT_BASE], xi[CT_BASE];
#pragma unroll
    for (int t = 0; t < CT_BASE; ++t) {
        const int r = row(t);
        T br = T(0), bi = T(0);
        if (r < nb) {
            const T *b = LEFT ? B + 2 * (r + q * ldb) : B + 2 * (q + (long)r * ldb);
            br = b[0];
            bi = b[1];
        }
        xr[t] = ar * br - ai * bi;
        xi[t] = ar * bi + ai * br;
    }
    // right-looking: x_t is divided when every earlier update has reached it,
    // then subtracted from each unknown after it -- independent updates, so
    // the dependent path per step is the division and one update (each
    // unknown still takes its subtractions in solve order)
#pragma unroll
    for (int s = 0; s < CT_BASE; ++s) {
        const int t = CT_BASE - 1 - s, rt = row(t);
        const T dr = Sr[rt][rt], di = Si[rt][rt];  // s / d = s conj(d) / |d|^2 (1 on a unit diagonal)
        const T den = dr * dr + di * di;
        const T tr_ = (xr[t] * dr + xi[t] * di) / den, ti_ = (xi[t] * dr - xr[t] * di) / den;
        xr[t] = tr_;
        xi[t] = ti_;
#pragma unroll
        for (int u = 0; u < t; ++u) {
            const int v = t - 1 - u, rv = row(v);  // solved after t
            const T cr = LEFT ? Sr[rv][rt] : Sr[rt][rv], ci = LEFT ? Si[rv][rt] : Si[rt][rv];
            xr[v] -= cr * tr_ - ci * ti_;
            xi[v] -= cr * ti_ + ci * tr_;
            if ((u & 7) == 7) __builtin_amdgcn_sched_barrier(0);  // 8 coefficients at a time
        }
        __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int t = 0; t < CT_BASE; ++t) {
        const int r = row(t);
        if (r < nb) {
            T *b = LEFT ? B + 2 * (r + q * ldb) : B + 2 * (q + (long)r * ldb);
            b[0] = xr[t];
            b[1] = xi[t];
        }
    }
}

// left: op(A) X = alpha B with op(A) of order m (rows of B); right: X op(A)
// = alpha B with op(A) of order n (columns of B).  lower: op(A) is lower
// triangular.
template <typename T>
int ctrsm_rec(hipStream_t s, bool left, bool lower, char tr, bool unit, int m, int n, T ar, T ai,
              const T *A, int lda, T *B, int ldb)
{
    const int na = left ? m : n;
    const int trc = (tr == 'N' || tr == 'n') ? 0 : ((tr == 'T' || tr == 't') ? 1 : 2);
    if (na <= CT_BASE) {
        const int nrhs = left ? n : m;
        const dim3 grid((unsigned)((nrhs + CT_THREADS - 1) / CT_THREADS));
#define GB_CTB(L, F)                                                                             \
    hipLaunchKernelGGL((ctrsm_base<T, L, F>), grid, dim3(CT_THREADS), 0, s, na, nrhs, trc,       \
                       unit ? 1 : 0, ar, ai, A, (long)lda, B, (long)ldb)
        // forward order for a lower op(A) on the left and an upper one on the right
        PROF_BEGIN(s, sizeof(T) == 8 ? "ztrsm_base" : "ctrsm_base");
        if (left) {
            if (lower) GB_CTB(true, true);
            else GB_CTB(true, false);
        } else {
            if (lower) GB_CTB(false, false);
            else GB_CTB(false, true);
        }
#undef GB_CTB
        PROF_END(s);
        return (int)hipGetLastError();
    }
    const int n1 = na / 2 / 16 * 16 > 0 ? na / 2 / 16 * 16 : na / 2, n2 = na - n1;
    // op(A) blocks [[A11, A12], [A21, A22]]: stored block (r0, c0) of op(A)'s
    // block (R, C) is (R, C) for 'N' and (C, R) otherwise
    auto blk = [&](int R, int C) -> const T * {
        const long r0 = R ? n1 : 0, c0 = C ? n1 : 0;
        return trc == 0 ? A + 2 * (r0 + c0 * lda) : A + 2 * (c0 + r0 * lda);
    };
    int err;
    if (left) {
        T *B1 = B, *B2 = B + 2 * (long)n1;
        if (lower) {  // X1 = A11 \ aB1; B2 = aB2 - A21 X1; X2 = A22 \ B2
            err = ctrsm_rec<T>(s, true, true, tr, unit, n1, n, ar, ai, blk(0, 0), lda, B1, ldb);
            if (!err) err = cgemm_entry<T>(s, tr, 'N', n2, n, n1, T(-1), T(0), blk(1, 0), lda, B1, ldb, ar, ai, B2, ldb);
            if (!err) err = ctrsm_rec<T>(s, true, true, tr, unit, n2, n, T(1), T(0), blk(1, 1), lda, B2, ldb);
        } else {      // X2 = A22 \ aB2; B1 = aB1 - A12 X2; X1 = A11 \ B1
            err = ctrsm_rec<T>(s, true, false, tr, unit, n2, n, ar, ai, blk(1, 1), lda, B2, ldb);
            if (!err) err = cgemm_entry<T>(s, tr, 'N', n1, n, n2, T(-1), T(0), blk(0, 1), lda, B2, ldb, ar, ai, B1, ldb);
            if (!err) err = ctrsm_rec<T>(s, true, false, tr, unit, n1, n, T(1), T(0), blk(0, 0), lda, B1, ldb);
        }
    } else {
        T *B1 = B, *B2 = B + 2 * (long)n1 * ldb;
        if (lower) {  // X2 = aB2 / A22; B1 = aB1 - X2 A21; X1 = B1 / A11
            err = ctrsm_rec<T>(s, false, true, tr, unit, m, n2, ar, ai, blk(1, 1), lda, B2, ldb);
            if (!err) err = cgemm_entry<T>(s, 'N', tr, m, n1, n2, T(-1), T(0), B2, ldb, blk(1, 0), lda, ar, ai, B1, ldb);
            if (!err) err = ctrsm_rec<T>(s, false, true, tr, unit, m, n1, T(1), T(0), blk(0, 0), lda, B1, ldb);
        } else {      // X1 = aB1 / A11; B2 = aB2 - X1 A12; X2 = B2 / A22
            err = ctrsm_rec<T>(s, false, false, tr, unit, m, n1, ar, ai, blk(0, 0), lda, B1, ldb);
            if (!err) err = cgemm_entry<T>(s, 'N', tr, m, n2, n1, T(-1), T(0), B1, ldb, blk(0, 1), lda, ar, ai, B2, ldb);
            if (!err) err = ctrsm_rec<T>(s, false, false, tr, unit, m, n2, T(1), T(0), blk(1, 1), lda, B2, ldb);
        }
    }
    return err;
}

template <typename T>
int ctrsm_entry(void *stream, char side, char uplo, char transa, char diag, int m, int n, T ar,
                T ai, const T *A, int lda, T *B, int ldb)
{
    hipStream_t s = pick_stream(stream);
    if (m <= 0 || n <= 0) return 0;
    if (ar == T(0) && ai == T(0))
        return cgemm_entry<T>(s, 'N', 'N', m, n, 0, T(0), T(0), B, ldb, B, ldb, T(0), T(0), B, ldb);
    const bool left = side == 'L' || side == 'l';
    const bool up = uplo == 'U' || uplo == 'u', nt = is_n(transa);
    const bool lower = up != nt;  // op(A) lower triangular
    return ctrsm_rec<T>(s, left, lower, transa, diag == 'U' || diag == 'u', m, n, ar, ai, A, lda,
                        B, ldb);
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

int gpublas_dev_zsyrk(void *stream, char uplo, char trans, int n, int k, double alpha_re,
                      double alpha_im, const double *d_A, int lda, double beta_re, double beta_im,
                      double *d_C, int ldc)
{
    return crk_entry<double>(stream, K_SYRK, uplo, trans, n, k, alpha_re, alpha_im, d_A, lda,
                             nullptr, 0, beta_re, beta_im, d_C, ldc);
}
int gpublas_dev_csyrk(void *stream, char uplo, char trans, int n, int k, float alpha_re,
                      float alpha_im, const float *d_A, int lda, float beta_re, float beta_im,
                      float *d_C, int ldc)
{
    return crk_entry<float>(stream, K_SYRK, uplo, trans, n, k, alpha_re, alpha_im, d_A, lda,
                            nullptr, 0, beta_re, beta_im, d_C, ldc);
}
int gpublas_dev_zherk(void *stream, char uplo, char trans, int n, int k, double alpha,
                      const double *d_A, int lda, double beta, double *d_C, int ldc)
{
    return crk_entry<double>(stream, K_HERK, uplo, trans, n, k, alpha, 0.0, d_A, lda, nullptr, 0,
                             beta, 0.0, d_C, ldc);
}
int gpublas_dev_cherk(void *stream, char uplo, char trans, int n, int k, float alpha,
                      const float *d_A, int lda, float beta, float *d_C, int ldc)
{
    return crk_entry<float>(stream, K_HERK, uplo, trans, n, k, alpha, 0.0f, d_A, lda, nullptr, 0,
                            beta, 0.0f, d_C, ldc);
}
int gpublas_dev_zsyr2k(void *stream, char uplo, char trans, int n, int k, double alpha_re,
                       double alpha_im, const double *d_A, int lda, const double *d_B, int ldb,
                       double beta_re, double beta_im, double *d_C, int ldc)
{
    return crk_entry<double>(stream, K_SYR2K, uplo, trans, n, k, alpha_re, alpha_im, d_A, lda,
                             d_B, ldb, beta_re, beta_im, d_C, ldc);
}
int gpublas_dev_csyr2k(void *stream, char uplo, char trans, int n, int k, float alpha_re,
                       float alpha_im, const float *d_A, int lda, const float *d_B, int ldb,
                       float beta_re, float beta_im, float *d_C, int ldc)
{
    return crk_entry<float>(stream, K_SYR2K, uplo, trans, n, k, alpha_re, alpha_im, d_A, lda, d_B,
                            ldb, beta_re, beta_im, d_C, ldc);
}
int gpublas_dev_zher2k(void *stream, char uplo, char trans, int n, int k, double alpha_re,
                       double alpha_im, const double *d_A, int lda, const double *d_B, int ldb,
                       double beta, double *d_C, int ldc)
{
    return crk_entry<double>(stream, K_HER2K, uplo, trans, n, k, alpha_re, alpha_im, d_A, lda,
                             d_B, ldb, beta, 0.0, d_C, ldc);
}
int gpublas_dev_cher2k(void *stream, char uplo, char trans, int n, int k, float alpha_re,
                       float alpha_im, const float *d_A, int lda, const float *d_B, int ldb,
                       float beta, float *d_C, int ldc)
{
    return crk_entry<float>(stream, K_HER2K, uplo, trans, n, k, alpha_re, alpha_im, d_A, lda, d_B,
                            ldb, beta, 0.0f, d_C, ldc);
}
int gpublas_dev_zsymm(void *stream, int herm, char side, char uplo, int m, int n,
                      double alpha_re, double alpha_im, const double *d_A, int lda,
                      const double *d_B, int ldb, double beta_re, double beta_im, double *d_C,
                      int ldc)
{
    return csymm_entry<double>(stream, herm, side, uplo, m, n, alpha_re, alpha_im, d_A, lda, d_B,
                               ldb, beta_re, beta_im, d_C, ldc);
}
int gpublas_dev_csymm(void *stream, int herm, char side, char uplo, int m, int n, float alpha_re,
                      float alpha_im, const float *d_A, int lda, const float *d_B, int ldb,
                      float beta_re, float beta_im, float *d_C, int ldc)
{
    return csymm_entry<float>(stream, herm, side, uplo, m, n, alpha_re, alpha_im, d_A, lda, d_B,
                              ldb, beta_re, beta_im, d_C, ldc);
}
int gpublas_dev_ztrmm(void *stream, char side, char uplo, char transa, char diag, int m, int n,
                      double alpha_re, double alpha_im, const double *d_A, int lda, double *d_B,
                      int ldb)
{
    return ctrmm_entry<double>(stream, side, uplo, transa, diag, m, n, alpha_re, alpha_im, d_A,
                               lda, d_B, ldb);
}
int gpublas_dev_ctrmm(void *stream, char side, char uplo, char transa, char diag, int m, int n,
                      float alpha_re, float alpha_im, const float *d_A, int lda, float *d_B,
                      int ldb)
{
    return ctrmm_entry<float>(stream, side, uplo, transa, diag, m, n, alpha_re, alpha_im, d_A, lda,
                              d_B, ldb);
}
int gpublas_dev_ztrsm(void *stream, char side, char uplo, char transa, char diag, int m, int n,
                      double alpha_re, double alpha_im, const double *d_A, int lda, double *d_B,
                      int ldb)
{
    return ctrsm_entry<double>(stream, side, uplo, transa, diag, m, n, alpha_re, alpha_im, d_A,
                               lda, d_B, ldb);
}
int gpublas_dev_ctrsm(void *stream, char side, char uplo, char transa, char diag, int m, int n,
                      float alpha_re, float alpha_im, const float *d_A, int lda, float *d_B,
                      int ldb)
{
    return ctrsm_entry<float>(stream, side, uplo, transa, diag, m, n, alpha_re, alpha_im, d_A, lda,
                              d_B, ldb);
}

}  // extern "C"
