/*
 * oracle/refblas.c -- CPU restatement of the reference's BLAS semantics for
 * the interposed hot path.  TEST INFRASTRUCTURE ONLY (see refblas.h).
 *
 * Argument checks follow the reference's netlib-style check macros:
 *   gemm  gemm.cc:87-127   (info 1,2,3,4,5,8,10,13)
 *   syrk  syrk.cc:78-146   (info 1,2,3,4,7,10)
 *   trsm  trsm.cc:76-132   (info 1,2,3,4,5,6,9,11)
 *   symm  symm.cc:81-122   (info 1,2,3,4,7,9,12)
 *   syr2k syr2k.cc:79-114  (info 1,2,3,4,7,9,12)
 *   trmm  netlib DTRMM     (info 1,2,3,4,5,6,9,11; trmm.cc is not built,
 *                           meson.build:87)
 * and, for routines the reference only specifies through its unbuilt CBLAS
 * wrappers (blas_level1/axpy.cc, dot.cc, nrm2.cc, amax.cc,
 * blas_level2/gemv.cc, ger.cc, trsv.cc), the netlib reference BLAS
 * definition those wrappers forward to.  Quick returns follow netlib, not the
 * reference's known defects (gemm.cc:121-126 multiplies C by beta==0, which
 * propagates NaN; syrk.cc:122-145 scales C twice): netlib sets C = 0 when
 * beta == 0 and returns after the alpha == 0 scaling.
 *
 * Arithmetic follows the netlib loop order (column sweeps with axpy inner
 * loops for the 'N' forms, dot-product inner loops for the 'T' forms), which
 * is what the fall-through host BLAS computes up to summation order.
 */
#include "refblas.h"
#include <complex.h>
#include <stdlib.h>
#include <math.h>
#include <stddef.h>
#include <ctype.h>

int ora_lsame(char a, char b) { return toupper((unsigned char)a) == toupper((unsigned char)b); }

static int imax(int a, int b) { return a > b ? a : b; }

#define IX(i, j, ld) ((size_t)(j) * (size_t)(ld) + (size_t)(i))

/* ------------------------------------------------------------------ */
/* INFO codes                                                          */
/* ------------------------------------------------------------------ */
int ora_gemm_info(char ta, char tb, int m, int n, int k, int lda, int ldb, int ldc)
{
    /* gemm.cc:87-116 */
    int nota = ora_lsame(ta, 'N'), notb = ora_lsame(tb, 'N');
    int nrowa = nota ? m : k, nrowb = notb ? k : n;
    if (!nota && !ora_lsame(ta, 'C') && !ora_lsame(ta, 'T')) return 1;
    if (!notb && !ora_lsame(tb, 'C') && !ora_lsame(tb, 'T')) return 2;
    if (m < 0) return 3;
    if (n < 0) return 4;
    if (k < 0) return 5;
    if (lda < imax(1, nrowa)) return 8;
    if (ldb < imax(1, nrowb)) return 10;
    if (ldc < imax(1, m)) return 13;
    return 0;
}

int ora_syrk_info(char uplo, char trans, int n, int k, int lda, int ldc)
{
    /* syrk.cc:87-115 */
    int nrowa = ora_lsame(trans, 'N') ? n : k;
    int upper = ora_lsame(uplo, 'U');
    if (!upper && !ora_lsame(uplo, 'L')) return 1;
    if (!ora_lsame(trans, 'N') && !ora_lsame(trans, 'T') && !ora_lsame(trans, 'C')) return 2;
    if (n < 0) return 3;
    if (k < 0) return 4;
    if (lda < imax(1, nrowa)) return 7;
    if (ldc < imax(1, n)) return 10;
    return 0;
}

int ora_trsm_info(char side, char uplo, char ta, char diag, int m, int n, int lda, int ldb)
{
    /* trsm.cc:84-116 */
    int lside = ora_lsame(side, 'L');
    int nrowa = lside ? m : n;
    int upper = ora_lsame(uplo, 'U');
    if (!lside && !ora_lsame(side, 'R')) return 1;
    if (!upper && !ora_lsame(uplo, 'L')) return 2;
    if (!ora_lsame(ta, 'N') && !ora_lsame(ta, 'T') && !ora_lsame(ta, 'C')) return 3;
    if (!ora_lsame(diag, 'U') && !ora_lsame(diag, 'N')) return 4;
    if (m < 0) return 5;
    if (n < 0) return 6;
    if (lda < imax(1, nrowa)) return 9;
    if (ldb < imax(1, m)) return 11;
    return 0;
}

int ora_symm_info(char side, char uplo, int m, int n, int lda, int ldb, int ldc)
{
    /* symm.cc:81-104 */
    int nrowa = ora_lsame(side, 'L') ? m : n;
    int upper = ora_lsame(uplo, 'U');
    if (!ora_lsame(side, 'L') && !ora_lsame(side, 'R')) return 1;
    if (!upper && !ora_lsame(uplo, 'L')) return 2;
    if (m < 0) return 3;
    if (n < 0) return 4;
    if (lda < imax(1, nrowa)) return 7;
    if (ldb < imax(1, m)) return 9;
    if (ldc < imax(1, m)) return 12;
    return 0;
}

int ora_gemv_info(char trans, int m, int n, int lda, int incx, int incy)
{
    /* netlib DGEMV, the routine cblas_dgemv forwards to (gemv.cc:11-98) */
    if (!ora_lsame(trans, 'N') && !ora_lsame(trans, 'T') && !ora_lsame(trans, 'C')) return 1;
    if (m < 0) return 2;
    if (n < 0) return 3;
    if (lda < imax(1, m)) return 6;
    if (incx == 0) return 8;
    if (incy == 0) return 11;
    return 0;
}

int ora_ger_info(int m, int n, int incx, int incy, int lda)
{
    /* netlib DGER (ger.cc:11-94) */
    if (m < 0) return 1;
    if (n < 0) return 2;
    if (incx == 0) return 5;
    if (incy == 0) return 7;
    if (lda < imax(1, m)) return 9;
    return 0;
}

int ora_trsv_info(char uplo, char trans, char diag, int n, int lda, int incx)
{
    /* netlib DTRSV (trsv.cc:11-88) */
    if (!ora_lsame(uplo, 'U') && !ora_lsame(uplo, 'L')) return 1;
    if (!ora_lsame(trans, 'N') && !ora_lsame(trans, 'T') && !ora_lsame(trans, 'C')) return 2;
    if (!ora_lsame(diag, 'U') && !ora_lsame(diag, 'N')) return 3;
    if (n < 0) return 4;
    if (lda < imax(1, n)) return 6;
    if (incx == 0) return 8;
    return 0;
}

int ora_symv_info(char uplo, int n, int lda, int incx, int incy)
{
    /* netlib DSYMV, which blas_level2/symv.cc:11-75 forwards to */
    if (!ora_lsame(uplo, 'U') && !ora_lsame(uplo, 'L')) return 1;
    if (n < 0) return 2;
    if (lda < imax(1, n)) return 5;
    if (incx == 0) return 7;
    if (incy == 0) return 10;
    return 0;
}

int ora_syr2_info(char uplo, int n, int incx, int incy, int lda)
{
    /* netlib DSYR2, which blas_level2/syr2.cc forwards to */
    if (!ora_lsame(uplo, 'U') && !ora_lsame(uplo, 'L')) return 1;
    if (n < 0) return 2;
    if (incx == 0) return 5;
    if (incy == 0) return 7;
    if (lda < imax(1, n)) return 9;
    return 0;
}

int ora_spr_info(char uplo, int n, int incx)
{
    /* netlib DSPR, which blas_level2/spr.cc forwards to */
    if (!ora_lsame(uplo, 'U') && !ora_lsame(uplo, 'L')) return 1;
    if (n < 0) return 2;
    if (incx == 0) return 5;
    return 0;
}

int ora_spr2_info(char uplo, int n, int incx, int incy)
{
    /* netlib DSPR2, which blas_level2/spr2.cc forwards to */
    if (!ora_lsame(uplo, 'U') && !ora_lsame(uplo, 'L')) return 1;
    if (n < 0) return 2;
    if (incx == 0) return 5;
    if (incy == 0) return 7;
    return 0;
}

int ora_trmv_info(char uplo, char trans, char diag, int n, int lda, int incx)
{
    /* netlib DTRMV, which blas_level2/trmv.cc forwards to */
    if (!ora_lsame(uplo, 'U') && !ora_lsame(uplo, 'L')) return 1;
    if (!ora_lsame(trans, 'N') && !ora_lsame(trans, 'T') && !ora_lsame(trans, 'C')) return 2;
    if (!ora_lsame(diag, 'U') && !ora_lsame(diag, 'N')) return 3;
    if (n < 0) return 4;
    if (lda < imax(1, n)) return 6;
    if (incx == 0) return 8;
    return 0;
}

int ora_tpmv_info(char uplo, char trans, char diag, int n, int incx)
{
    /* netlib DTPMV / DTPSV, which blas_level2/tpmv.cc and tpsv.cc forward to */
    if (!ora_lsame(uplo, 'U') && !ora_lsame(uplo, 'L')) return 1;
    if (!ora_lsame(trans, 'N') && !ora_lsame(trans, 'T') && !ora_lsame(trans, 'C')) return 2;
    if (!ora_lsame(diag, 'U') && !ora_lsame(diag, 'N')) return 3;
    if (n < 0) return 4;
    if (incx == 0) return 7;
    return 0;
}

int ora_spmv_info(char uplo, int n, int incx, int incy)
{
    /* netlib DSPMV, which blas_level2/pmv.cc forwards to */
    if (!ora_lsame(uplo, 'U') && !ora_lsame(uplo, 'L')) return 1;
    if (n < 0) return 2;
    if (incx == 0) return 6;
    if (incy == 0) return 9;
    return 0;
}

int ora_syr_info(char uplo, int n, int incx, int lda)
{
    /* netlib DSYR, which blas_level2/syr.cc:11-68 forwards to */
    if (!ora_lsame(uplo, 'U') && !ora_lsame(uplo, 'L')) return 1;
    if (n < 0) return 2;
    if (incx == 0) return 5;
    if (lda < imax(1, n)) return 7;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Level 3, generated for double and float                            */
/* ------------------------------------------------------------------ */
#define DEFINE_GEMM(NAME, T)                                                                  \
void NAME(char ta, char tb, int m, int n, int k, T alpha, const T *A, int lda, const T *B,    \
          int ldb, T beta, T *C, int ldc)                                                     \
{                                                                                             \
    int nota = ora_lsame(ta, 'N'), notb = ora_lsame(tb, 'N');                                 \
    if (ora_gemm_info(ta, tb, m, n, k, lda, ldb, ldc)) return;                                \
    if (m == 0 || n == 0 || ((alpha == 0 || k == 0) && beta == 1)) return;                    \
    if (alpha == 0) {                                                                         \
        for (int j = 0; j < n; ++j)                                                           \
            for (int i = 0; i < m; ++i)                                                       \
                C[IX(i, j, ldc)] = beta == 0 ? (T)0 : beta * C[IX(i, j, ldc)];                \
        return;                                                                               \
    }                                                                                         \
    if (nota) {                                                                               \
        /* C := alpha*A*op(B) + beta*C, column axpy form */                                   \
        for (int j = 0; j < n; ++j) {                                                         \
            if (beta == 0)                                                                    \
                for (int i = 0; i < m; ++i) C[IX(i, j, ldc)] = 0;                             \
            else if (beta != 1)                                                               \
                for (int i = 0; i < m; ++i) C[IX(i, j, ldc)] *= beta;                         \
            for (int l = 0; l < k; ++l) {                                                     \
                T temp = alpha * (notb ? B[IX(l, j, ldb)] : B[IX(j, l, ldb)]);                \
                for (int i = 0; i < m; ++i) C[IX(i, j, ldc)] += temp * A[IX(i, l, lda)];      \
            }                                                                                 \
        }                                                                                     \
    } else {                                                                                  \
        /* C := alpha*A**T*op(B) + beta*C, dot form */                                        \
        for (int j = 0; j < n; ++j)                                                           \
            for (int i = 0; i < m; ++i) {                                                     \
                T temp = 0;                                                                   \
                for (int l = 0; l < k; ++l)                                                   \
                    temp += A[IX(l, i, lda)] * (notb ? B[IX(l, j, ldb)] : B[IX(j, l, ldb)]);  \
                C[IX(i, j, ldc)] = beta == 0 ? alpha * temp                                   \
                                             : alpha * temp + beta * C[IX(i, j, ldc)];        \
            }                                                                                 \
    }                                                                                         \
}
DEFINE_GEMM(ora_dgemm, double)
DEFINE_GEMM(ora_sgemm, float)

/* Complex GEMM, netlib ZGEMM / CGEMM (the routines behind the reference's
 * F77_gemm(z) / F77_gemm(c), gemm.cc:181-217): 'C' conjugates; the 'N' form
 * of A runs column axpys, the others dot products, as netlib does.  Operands
 * and scalars are (re, im) pairs. */
#define DEFINE_CGEMM(NAME, R, CT)                                                             \
void NAME(char ta, char tb, int m, int n, int k, const R *alpha_, const R *A_, int lda,       \
          const R *B_, int ldb, const R *beta_, R *C_, int ldc)                               \
{                                                                                             \
    const CT *A = (const CT *)A_, *B = (const CT *)B_;                                        \
    CT *C = (CT *)C_;                                                                         \
    const CT alpha = *(const CT *)alpha_, beta = *(const CT *)beta_;                          \
    int nota = ora_lsame(ta, 'N'), notb = ora_lsame(tb, 'N');                                 \
    int conja = ora_lsame(ta, 'C'), conjb = ora_lsame(tb, 'C');                               \
    if (ora_gemm_info(ta, tb, m, n, k, lda, ldb, ldc)) return;                                \
    if (m == 0 || n == 0 || ((alpha == 0 || k == 0) && beta == 1)) return;                    \
    if (alpha == 0) {                                                                         \
        for (int j = 0; j < n; ++j)                                                           \
            for (int i = 0; i < m; ++i)                                                       \
                C[IX(i, j, ldc)] = beta == 0 ? (CT)0 : beta * C[IX(i, j, ldc)];               \
        return;                                                                               \
    }                                                                                         \
    for (int j = 0; j < n; ++j) {                                                             \
        if (nota) {                                                                           \
            if (beta == 0)                                                                    \
                for (int i = 0; i < m; ++i) C[IX(i, j, ldc)] = 0;                             \
            else if (beta != 1)                                                               \
                for (int i = 0; i < m; ++i) C[IX(i, j, ldc)] *= beta;                         \
            for (int l = 0; l < k; ++l) {                                                     \
                CT b = notb ? B[IX(l, j, ldb)] : B[IX(j, l, ldb)];                            \
                if (conjb) b = conj(b);                                                       \
                CT temp = alpha * b;                                                          \
                for (int i = 0; i < m; ++i) C[IX(i, j, ldc)] += temp * A[IX(i, l, lda)];      \
            }                                                                                 \
        } else {                                                                              \
            for (int i = 0; i < m; ++i) {                                                     \
                CT temp = 0;                                                                  \
                for (int l = 0; l < k; ++l) {                                                 \
                    CT a = A[IX(l, i, lda)];                                                  \
                    CT b = notb ? B[IX(l, j, ldb)] : B[IX(j, l, ldb)];                        \
                    if (conja) a = conj(a);                                                   \
                    if (conjb) b = conj(b);                                                   \
                    temp += a * b;                                                            \
                }                                                                             \
                C[IX(i, j, ldc)] = beta == 0 ? alpha * temp                                   \
                                             : alpha * temp + beta * C[IX(i, j, ldc)];        \
            }                                                                                 \
        }                                                                                     \
    }                                                                                         \
}
DEFINE_CGEMM(ora_zgemm, double, double _Complex)
DEFINE_CGEMM(ora_cgemm, float, float _Complex)

/* The rest of the complex Level-3 (netlib ZSYRK / ZHERK / ZSYR2K / ZHER2K /
 * ZSYMM / ZHEMM / ZTRMM / ZTRSM, the routines behind the reference's
 * syrk.cc:182-215, herk.cc:131-162, syr2k.cc:182-215, her2k.cc:141-178,
 * symm.cc:158-207, hemm.cc:120-153, trsm.cc:175-215; ZTRMM as netlib, trmm.cc
 * being unbuilt).  Each result element is its defining sum taken in index
 * order; argument errors return without touching the outputs (the library's
 * INFO codes are checked by the reference's own zblat3 / cblat3).  Scalars
 * and operands are (re, im) pairs; HERK's alpha / beta and HER2K's beta are
 * real. */
#define DEFINE_CRK(NAME, R, CT, HERM, TWO)                                                     \
void NAME(char uplo, char trans, int n, int k, const R *alpha_, const R *A_, int lda,         \
          const R *B_, int ldb, const R *beta_, R *C_, int ldc)                               \
{                                                                                             \
    const CT *A = (const CT *)A_, *B = (const CT *)B_;                                        \
    CT *C = (CT *)C_;                                                                         \
    const CT alpha = HERM && !TWO ? (CT)alpha_[0] : *(const CT *)alpha_;                      \
    const CT beta = HERM ? (CT)beta_[0] : *(const CT *)beta_;                                 \
    const int upper = ora_lsame(uplo, 'U'), nt = ora_lsame(trans, 'N');                      \
    if (n < 0 || k < 0) return;                                                               \
    if (n == 0 || ((alpha == 0 || k == 0) && beta == 1)) return;                              \
    for (int j = 0; j < n; ++j)                                                               \
        for (int i = upper ? 0 : j; i < (upper ? j + 1 : n); ++i) {                           \
            CT acc = 0;                                                                       \
            if (alpha != 0)                                                                   \
                for (int l = 0; l < k; ++l) {                                                 \
                    /* x = op(A)(i, l), xj = op(A)(j, l); y likewise for B */                 \
                    CT xi = nt ? A[IX(i, l, lda)] : A[IX(l, i, lda)];                         \
                    CT xj = nt ? A[IX(j, l, lda)] : A[IX(l, j, lda)];                         \
                    if (HERM && !nt) { xi = conj(xi); xj = conj(xj); }                        \
                    if (!TWO) {                                                               \
                        acc += xi * (HERM ? conj(xj) : xj);                                   \
                    } else {                                                                  \
                        CT yi = nt ? B[IX(i, l, ldb)] : B[IX(l, i, ldb)];                     \
                        CT yj = nt ? B[IX(j, l, ldb)] : B[IX(l, j, ldb)];                     \
                        if (HERM && !nt) { yi = conj(yi); yj = conj(yj); }                    \
                        acc += HERM ? alpha * xi * conj(yj) + conj(alpha) * yi * conj(xj)     \
                                    : alpha * (xi * yj + yi * xj);                            \
                    }                                                                         \
                }                                                                             \
            CT v = TWO ? acc : alpha * acc;                                                   \
            if (beta != 0) {                                                                  \
                CT c = C[IX(i, j, ldc)];                                                      \
                if (HERM && i == j) c = creal(c);                                             \
                v += beta * c;                                                                \
            }                                                                                 \
            C[IX(i, j, ldc)] = (HERM && i == j) ? (CT)creal(v) : v;                           \
        }                                                                                     \
}
DEFINE_CRK(ora_zsyrk_, double, double _Complex, 0, 0)
DEFINE_CRK(ora_zherk_, double, double _Complex, 1, 0)
DEFINE_CRK(ora_zsyr2k, double, double _Complex, 0, 1)
DEFINE_CRK(ora_zher2k, double, double _Complex, 1, 1)
DEFINE_CRK(ora_csyrk_, float, float _Complex, 0, 0)
DEFINE_CRK(ora_cherk_, float, float _Complex, 1, 0)
DEFINE_CRK(ora_csyr2k, float, float _Complex, 0, 1)
DEFINE_CRK(ora_cher2k, float, float _Complex, 1, 1)

#define DEFINE_CSYK_WRAP(NAME, INNER, R)                                                      \
void NAME(char uplo, char trans, int n, int k, const R *alpha, const R *A, int lda,           \
          const R *beta, R *C, int ldc)                                                       \
{                                                                                             \
    INNER(uplo, trans, n, k, alpha, A, lda, A, lda, beta, C, ldc);                            \
}
DEFINE_CSYK_WRAP(ora_zsyrk, ora_zsyrk_, double)
DEFINE_CSYK_WRAP(ora_zherk, ora_zherk_, double)
DEFINE_CSYK_WRAP(ora_csyrk, ora_csyrk_, float)
DEFINE_CSYK_WRAP(ora_cherk, ora_cherk_, float)

/* element (i, j) of the full matrix a SYMM / HEMM reads from the uplo triangle */
#define CSYM_ELEM(A, i, j, lda, upper, herm)                                                  \
    (((upper) ? (i) <= (j) : (i) >= (j))                                                      \
         ? (((herm) && (i) == (j)) ? creal(A[IX(i, j, lda)]) : A[IX(i, j, lda)])              \
         : ((herm) ? conj(A[IX(j, i, lda)]) : A[IX(j, i, lda)]))

#define DEFINE_CSYMM(NAME, R, CT, HERM)                                                       \
void NAME(char side, char uplo, int m, int n, const R *alpha_, const R *A_, int lda,          \
          const R *B_, int ldb, const R *beta_, R *C_, int ldc)                               \
{                                                                                             \
    const CT *A = (const CT *)A_, *B = (const CT *)B_;                                        \
    CT *C = (CT *)C_;                                                                         \
    const CT alpha = *(const CT *)alpha_, beta = *(const CT *)beta_;                          \
    const int left = ora_lsame(side, 'L'), upper = ora_lsame(uplo, 'U');                      \
    if (m <= 0 || n <= 0 || (alpha == 0 && beta == 1)) return;                                \
    for (int j = 0; j < n; ++j)                                                               \
        for (int i = 0; i < m; ++i) {                                                         \
            CT acc = 0;                                                                       \
            if (alpha != 0) {                                                                 \
                if (left)                                                                     \
                    for (int l = 0; l < m; ++l)                                               \
                        acc += CSYM_ELEM(A, i, l, lda, upper, HERM) * B[IX(l, j, ldb)];       \
                else                                                                          \
                    for (int l = 0; l < n; ++l)                                               \
                        acc += B[IX(i, l, ldb)] * CSYM_ELEM(A, l, j, lda, upper, HERM);       \
            }                                                                                 \
            CT v = alpha * acc;                                                               \
            if (beta != 0) v += beta * C[IX(i, j, ldc)];                                      \
            C[IX(i, j, ldc)] = v;                                                             \
        }                                                                                     \
}
DEFINE_CSYMM(ora_zsymm, double, double _Complex, 0)
DEFINE_CSYMM(ora_zhemm, double, double _Complex, 1)
DEFINE_CSYMM(ora_csymm, float, float _Complex, 0)
DEFINE_CSYMM(ora_chemm, float, float _Complex, 1)

/* op(T)(r, c) of the triangle: zero outside it, 1 on a unit diagonal */
#define CTRI_OP(A, r, c, lda, upper, tr, unit)                                                \
    ({                                                                                        \
        const int r_ = (tr) ? (c) : (r), c_ = (tr) ? (r) : (c);                               \
        const int in_ = (upper) ? r_ <= c_ : r_ >= c_;                                        \
        __typeof__(A[0]) v_ = !in_ ? 0 : (r_ == c_ && (unit)) ? 1 : A[IX(r_, c_, lda)];       \
        (tr) == 2 ? conj(v_) : v_;                                                            \
    })

#define DEFINE_CTRMM(NAME, R, CT)                                                             \
void NAME(char side, char uplo, char transa, char diag, int m, int n, const R *alpha_,        \
          const R *A_, int lda, R *B_, int ldb)                                               \
{                                                                                             \
    const CT *A = (const CT *)A_;                                                             \
    CT *B = (CT *)B_;                                                                         \
    const CT alpha = *(const CT *)alpha_;                                                     \
    const int left = ora_lsame(side, 'L'), upper = ora_lsame(uplo, 'U');                      \
    const int tr = ora_lsame(transa, 'N') ? 0 : ora_lsame(transa, 'T') ? 1 : 2;               \
    const int unit = ora_lsame(diag, 'U');                                                    \
    if (m <= 0 || n <= 0) return;                                                             \
    const int na = left ? m : n;                                                              \
    CT *W = (CT *)malloc((size_t)m * n * sizeof(CT));                                         \
    if (!W) return;                                                                           \
    for (int j = 0; j < n; ++j)                                                               \
        for (int i = 0; i < m; ++i) {                                                         \
            CT acc = 0;                                                                       \
            if (alpha != 0)                                                                   \
                for (int l = 0; l < na; ++l)                                                  \
                    acc += left ? CTRI_OP(A, i, l, lda, upper, tr, unit) * B[IX(l, j, ldb)]   \
                                : B[IX(i, l, ldb)] * CTRI_OP(A, l, j, lda, upper, tr, unit);  \
            W[IX(i, j, m)] = alpha * acc;                                                     \
        }                                                                                     \
    for (int j = 0; j < n; ++j)                                                               \
        for (int i = 0; i < m; ++i) B[IX(i, j, ldb)] = W[IX(i, j, m)];                        \
    free(W);                                                                                  \
}
DEFINE_CTRMM(ora_ztrmm, double, double _Complex)
DEFINE_CTRMM(ora_ctrmm, float, float _Complex)

/* substitution in solve order: x_t = (alpha b_t - sum of the solved terms) / a_tt */
#define DEFINE_CTRSM(NAME, R, CT)                                                             \
void NAME(char side, char uplo, char transa, char diag, int m, int n, const R *alpha_,        \
          const R *A_, int lda, R *B_, int ldb)                                               \
{                                                                                             \
    const CT *A = (const CT *)A_;                                                             \
    CT *B = (CT *)B_;                                                                         \
    const CT alpha = *(const CT *)alpha_;                                                     \
    const int left = ora_lsame(side, 'L'), upper = ora_lsame(uplo, 'U');                      \
    const int tr = ora_lsame(transa, 'N') ? 0 : ora_lsame(transa, 'T') ? 1 : 2;               \
    const int unit = ora_lsame(diag, 'U');                                                    \
    if (m <= 0 || n <= 0) return;                                                             \
    const int na = left ? m : n, nrhs = left ? n : m;                                         \
    /* op(A) lower: forward on the left, backward on the right */                             \
    const int lower = upper == (tr != 0);                                                     \
    const int fwd = left ? lower : !lower;                                                    \
    for (int q = 0; q < nrhs; ++q)                                                            \
        for (int s = 0; s < na; ++s) {                                                        \
            const int t = fwd ? s : na - 1 - s;                                               \
            CT *x = left ? &B[IX(t, q, ldb)] : &B[IX(q, t, ldb)];                             \
            CT v = alpha * *x;                                                                \
            for (int u = 0; u < s; ++u) {                                                     \
                const int w = fwd ? u : na - 1 - u;                                           \
                const CT xw = left ? B[IX(w, q, ldb)] : B[IX(q, w, ldb)];                     \
                v -= (left ? CTRI_OP(A, t, w, lda, upper, tr, 0)                              \
                           : CTRI_OP(A, w, t, lda, upper, tr, 0)) * xw;                       \
            }                                                                                 \
            if (!unit) v /= CTRI_OP(A, t, t, lda, upper, tr, 0);                              \
            *x = v;                                                                           \
        }                                                                                     \
}
DEFINE_CTRSM(ora_ztrsm, double, double _Complex)
DEFINE_CTRSM(ora_ctrsm, float, float _Complex)

#define DEFINE_SYRK(NAME, T)                                                                  \
void NAME(char uplo, char trans, int n, int k, T alpha, const T *A, int lda, T beta, T *C,    \
          int ldc)                                                                            \
{                                                                                             \
    int upper = ora_lsame(uplo, 'U');                                                         \
    if (ora_syrk_info(uplo, trans, n, k, lda, ldc)) return;                                   \
    if (n == 0 || ((alpha == 0 || k == 0) && beta == 1)) return;                              \
    if (alpha == 0) {                                                                         \
        for (int j = 0; j < n; ++j) {                                                         \
            int i0 = upper ? 0 : j, i1 = upper ? j + 1 : n;                                   \
            for (int i = i0; i < i1; ++i)                                                     \
                C[IX(i, j, ldc)] = beta == 0 ? (T)0 : beta * C[IX(i, j, ldc)];                \
        }                                                                                     \
        return;                                                                               \
    }                                                                                         \
    if (ora_lsame(trans, 'N')) {                                                              \
        for (int j = 0; j < n; ++j) {                                                         \
            int i0 = upper ? 0 : j, i1 = upper ? j + 1 : n;                                   \
            if (beta == 0)                                                                    \
                for (int i = i0; i < i1; ++i) C[IX(i, j, ldc)] = 0;                           \
            else if (beta != 1)                                                               \
                for (int i = i0; i < i1; ++i) C[IX(i, j, ldc)] *= beta;                       \
            for (int l = 0; l < k; ++l) {                                                     \
                T temp = alpha * A[IX(j, l, lda)];                                            \
                for (int i = i0; i < i1; ++i) C[IX(i, j, ldc)] += temp * A[IX(i, l, lda)];    \
            }                                                                                 \
        }                                                                                     \
    } else {                                                                                  \
        for (int j = 0; j < n; ++j) {                                                         \
            int i0 = upper ? 0 : j, i1 = upper ? j + 1 : n;                                   \
            for (int i = i0; i < i1; ++i) {                                                   \
                T temp = 0;                                                                   \
                for (int l = 0; l < k; ++l) temp += A[IX(l, i, lda)] * A[IX(l, j, lda)];      \
                C[IX(i, j, ldc)] = beta == 0 ? alpha * temp                                   \
                                             : alpha * temp + beta * C[IX(i, j, ldc)];        \
            }                                                                                 \
        }                                                                                     \
    }                                                                                         \
}
DEFINE_SYRK(ora_dsyrk, double)
DEFINE_SYRK(ora_ssyrk, float)

#define DEFINE_TRSM(NAME, T)                                                                  \
void NAME(char side, char uplo, char ta, char diag, int m, int n, T alpha, const T *A,        \
          int lda, T *B, int ldb)                                                             \
{                                                                                             \
    int lside = ora_lsame(side, 'L'), upper = ora_lsame(uplo, 'U');                           \
    int nounit = ora_lsame(diag, 'N'), notr = ora_lsame(ta, 'N');                             \
    if (ora_trsm_info(side, uplo, ta, diag, m, n, lda, ldb)) return;                          \
    if (m == 0 || n == 0) return;                                                             \
    if (alpha == 0) {                                                                         \
        for (int j = 0; j < n; ++j)                                                           \
            for (int i = 0; i < m; ++i) B[IX(i, j, ldb)] = 0;                                 \
        return;                                                                               \
    }                                                                                         \
    if (lside) {                                                                              \
        if (notr) {                                                                           \
            /* B := alpha*inv(A)*B */                                                         \
            for (int j = 0; j < n; ++j) {                                                     \
                if (alpha != 1)                                                               \
                    for (int i = 0; i < m; ++i) B[IX(i, j, ldb)] *= alpha;                    \
                if (upper) {                                                                  \
                    for (int kk = m - 1; kk >= 0; --kk)                                       \
                        if (B[IX(kk, j, ldb)] != 0) {                                         \
                            if (nounit) B[IX(kk, j, ldb)] /= A[IX(kk, kk, lda)];              \
                            for (int i = 0; i < kk; ++i)                                      \
                                B[IX(i, j, ldb)] -= B[IX(kk, j, ldb)] * A[IX(i, kk, lda)];    \
                        }                                                                     \
                } else {                                                                      \
                    for (int kk = 0; kk < m; ++kk)                                            \
                        if (B[IX(kk, j, ldb)] != 0) {                                         \
                            if (nounit) B[IX(kk, j, ldb)] /= A[IX(kk, kk, lda)];              \
                            for (int i = kk + 1; i < m; ++i)                                  \
                                B[IX(i, j, ldb)] -= B[IX(kk, j, ldb)] * A[IX(i, kk, lda)];    \
                        }                                                                     \
                }                                                                             \
            }                                                                                 \
        } else {                                                                              \
            /* B := alpha*inv(A**T)*B */                                                      \
            for (int j = 0; j < n; ++j) {                                                     \
                if (upper) {                                                                  \
                    for (int i = 0; i < m; ++i) {                                             \
                        T temp = alpha * B[IX(i, j, ldb)];                                    \
                        for (int kk = 0; kk < i; ++kk)                                        \
                            temp -= A[IX(kk, i, lda)] * B[IX(kk, j, ldb)];                    \
                        if (nounit) temp /= A[IX(i, i, lda)];                                 \
                        B[IX(i, j, ldb)] = temp;                                              \
                    }                                                                         \
                } else {                                                                      \
                    for (int i = m - 1; i >= 0; --i) {                                        \
                        T temp = alpha * B[IX(i, j, ldb)];                                    \
                        for (int kk = i + 1; kk < m; ++kk)                                    \
                            temp -= A[IX(kk, i, lda)] * B[IX(kk, j, ldb)];                    \
                        if (nounit) temp /= A[IX(i, i, lda)];                                 \
                        B[IX(i, j, ldb)] = temp;                                              \
                    }                                                                         \
                }                                                                             \
            }                                                                                 \
        }                                                                                     \
    } else {                                                                                  \
        if (notr) {                                                                           \
            /* B := alpha*B*inv(A) */                                                         \
            if (upper) {                                                                      \
                for (int j = 0; j < n; ++j) {                                                 \
                    if (alpha != 1)                                                           \
                        for (int i = 0; i < m; ++i) B[IX(i, j, ldb)] *= alpha;                \
                    for (int kk = 0; kk < j; ++kk)                                            \
                        if (A[IX(kk, j, lda)] != 0)                                           \
                            for (int i = 0; i < m; ++i)                                       \
                                B[IX(i, j, ldb)] -= A[IX(kk, j, lda)] * B[IX(i, kk, ldb)];    \
                    if (nounit) {                                                             \
                        T temp = (T)1 / A[IX(j, j, lda)];                                     \
                        for (int i = 0; i < m; ++i) B[IX(i, j, ldb)] *= temp;                 \
                    }                                                                         \
                }                                                                             \
            } else {                                                                          \
                for (int j = n - 1; j >= 0; --j) {                                            \
                    if (alpha != 1)                                                           \
                        for (int i = 0; i < m; ++i) B[IX(i, j, ldb)] *= alpha;                \
                    for (int kk = j + 1; kk < n; ++kk)                                        \
                        if (A[IX(kk, j, lda)] != 0)                                           \
                            for (int i = 0; i < m; ++i)                                       \
                                B[IX(i, j, ldb)] -= A[IX(kk, j, lda)] * B[IX(i, kk, ldb)];    \
                    if (nounit) {                                                             \
                        T temp = (T)1 / A[IX(j, j, lda)];                                     \
                        for (int i = 0; i < m; ++i) B[IX(i, j, ldb)] *= temp;                 \
                    }                                                                         \
                }                                                                             \
            }                                                                                 \
        } else {                                                                              \
            /* B := alpha*B*inv(A**T) */                                                      \
            if (upper) {                                                                      \
                for (int kk = n - 1; kk >= 0; --kk) {                                         \
                    if (nounit) {                                                             \
                        T temp = (T)1 / A[IX(kk, kk, lda)];                                   \
                        for (int i = 0; i < m; ++i) B[IX(i, kk, ldb)] *= temp;                \
                    }                                                                         \
                    for (int j = 0; j < kk; ++j)                                              \
                        if (A[IX(j, kk, lda)] != 0) {                                         \
                            T temp = A[IX(j, kk, lda)];                                       \
                            for (int i = 0; i < m; ++i)                                       \
                                B[IX(i, j, ldb)] -= temp * B[IX(i, kk, ldb)];                 \
                        }                                                                     \
                    if (alpha != 1)                                                           \
                        for (int i = 0; i < m; ++i) B[IX(i, kk, ldb)] *= alpha;               \
                }                                                                             \
            } else {                                                                          \
                for (int kk = 0; kk < n; ++kk) {                                              \
                    if (nounit) {                                                             \
                        T temp = (T)1 / A[IX(kk, kk, lda)];                                   \
                        for (int i = 0; i < m; ++i) B[IX(i, kk, ldb)] *= temp;                \
                    }                                                                         \
                    for (int j = kk + 1; j < n; ++j)                                          \
                        if (A[IX(j, kk, lda)] != 0) {                                         \
                            T temp = A[IX(j, kk, lda)];                                       \
                            for (int i = 0; i < m; ++i)                                       \
                                B[IX(i, j, ldb)] -= temp * B[IX(i, kk, ldb)];                 \
                        }                                                                     \
                    if (alpha != 1)                                                           \
                        for (int i = 0; i < m; ++i) B[IX(i, kk, ldb)] *= alpha;               \
                }                                                                             \
            }                                                                                 \
        }                                                                                     \
    }                                                                                         \
}
DEFINE_TRSM(ora_dtrsm, double)
DEFINE_TRSM(ora_strsm, float)

#define DEFINE_SYMM(NAME, T)                                                                  \
void NAME(char side, char uplo, int m, int n, T alpha, const T *A, int lda, const T *B,       \
          int ldb, T beta, T *C, int ldc)                                                     \
{                                                                                             \
    int upper = ora_lsame(uplo, 'U');                                                         \
    if (ora_symm_info(side, uplo, m, n, lda, ldb, ldc)) return;                               \
    if (m == 0 || n == 0 || (alpha == 0 && beta == 1)) return;                                \
    if (alpha == 0) {                                                                         \
        for (int j = 0; j < n; ++j)                                                           \
            for (int i = 0; i < m; ++i)                                                       \
                C[IX(i, j, ldc)] = beta == 0 ? (T)0 : beta * C[IX(i, j, ldc)];                \
        return;                                                                               \
    }                                                                                         \
    if (ora_lsame(side, 'L')) {                                                               \
        /* C := alpha*A*B + beta*C, A symmetric m x m */                                      \
        for (int j = 0; j < n; ++j) {                                                         \
            if (upper) {                                                                      \
                for (int i = 0; i < m; ++i) {                                                 \
                    T temp1 = alpha * B[IX(i, j, ldb)], temp2 = 0;                            \
                    for (int kk = 0; kk < i; ++kk) {                                          \
                        C[IX(kk, j, ldc)] += temp1 * A[IX(kk, i, lda)];                       \
                        temp2 += B[IX(kk, j, ldb)] * A[IX(kk, i, lda)];                       \
                    }                                                                         \
                    C[IX(i, j, ldc)] = (beta == 0 ? (T)0 : beta * C[IX(i, j, ldc)]) +         \
                                       temp1 * A[IX(i, i, lda)] + alpha * temp2;              \
                }                                                                             \
            } else {                                                                          \
                for (int i = m - 1; i >= 0; --i) {                                            \
                    T temp1 = alpha * B[IX(i, j, ldb)], temp2 = 0;                            \
                    for (int kk = i + 1; kk < m; ++kk) {                                      \
                        C[IX(kk, j, ldc)] += temp1 * A[IX(kk, i, lda)];                       \
                        temp2 += B[IX(kk, j, ldb)] * A[IX(kk, i, lda)];                       \
                    }                                                                         \
                    C[IX(i, j, ldc)] = (beta == 0 ? (T)0 : beta * C[IX(i, j, ldc)]) +         \
                                       temp1 * A[IX(i, i, lda)] + alpha * temp2;              \
                }                                                                             \
            }                                                                                 \
        }                                                                                     \
    } else {                                                                                  \
        /* C := alpha*B*A + beta*C, A symmetric n x n */                                      \
        for (int j = 0; j < n; ++j) {                                                         \
            T temp1 = alpha * A[IX(j, j, lda)];                                               \
            for (int i = 0; i < m; ++i)                                                       \
                C[IX(i, j, ldc)] = (beta == 0 ? (T)0 : beta * C[IX(i, j, ldc)]) +             \
                                   temp1 * B[IX(i, j, ldb)];                                  \
            for (int kk = 0; kk < n; ++kk) {                                                  \
                if (kk == j) continue;                                                        \
                T akj = (kk < j) == (upper != 0) ? A[IX(kk, j, lda)] : A[IX(j, kk, lda)];     \
                temp1 = alpha * akj;                                                          \
                for (int i = 0; i < m; ++i) C[IX(i, j, ldc)] += temp1 * B[IX(i, kk, ldb)];    \
            }                                                                                 \
        }                                                                                     \
    }                                                                                         \
}
DEFINE_SYMM(ora_dsymm, double)
DEFINE_SYMM(ora_ssymm, float)

/* ------------------------------------------------------------------ */
/* DSYR2K: info per syr2k.cc:79-114 (1 uplo, 2 trans, 3 n, 4 k, 7 lda,  */
/* 9 ldb, 12 ldc); arithmetic in the netlib DSYR2K loop order          */
/* ------------------------------------------------------------------ */
int ora_syr2k_info(char uplo, char trans, int n, int k, int lda, int ldb, int ldc)
{
    int nrowa = ora_lsame(trans, 'N') ? n : k;
    int upper = ora_lsame(uplo, 'U');
    if (!upper && !ora_lsame(uplo, 'L')) return 1;
    if (!ora_lsame(trans, 'N') && !ora_lsame(trans, 'T') && !ora_lsame(trans, 'C')) return 2;
    if (n < 0) return 3;
    if (k < 0) return 4;
    if (lda < imax(1, nrowa)) return 7;
    if (ldb < imax(1, nrowa)) return 9;
    if (ldc < imax(1, n)) return 12;
    return 0;
}

void ora_dsyr2k(char uplo, char trans, int n, int k, double alpha, const double *A, int lda,
                const double *B, int ldb, double beta, double *C, int ldc)
{
    int upper = ora_lsame(uplo, 'U');
    if (ora_syr2k_info(uplo, trans, n, k, lda, ldb, ldc)) return;
    if (n == 0 || ((alpha == 0 || k == 0) && beta == 1)) return;
    if (alpha == 0) {
        for (int j = 0; j < n; ++j) {
            int i0 = upper ? 0 : j, i1 = upper ? j + 1 : n;
            for (int i = i0; i < i1; ++i) C[IX(i, j, ldc)] = beta == 0 ? 0 : beta * C[IX(i, j, ldc)];
        }
        return;
    }
    if (ora_lsame(trans, 'N')) {
        /* C := alpha*A*B**T + alpha*B*A**T + beta*C, one column at a time */
        for (int j = 0; j < n; ++j) {
            int i0 = upper ? 0 : j, i1 = upper ? j + 1 : n;
            if (beta == 0)
                for (int i = i0; i < i1; ++i) C[IX(i, j, ldc)] = 0;
            else if (beta != 1)
                for (int i = i0; i < i1; ++i) C[IX(i, j, ldc)] *= beta;
            for (int l = 0; l < k; ++l) {
                if (A[IX(j, l, lda)] == 0 && B[IX(j, l, ldb)] == 0) continue;
                double t1 = alpha * B[IX(j, l, ldb)], t2 = alpha * A[IX(j, l, lda)];
                for (int i = i0; i < i1; ++i)
                    C[IX(i, j, ldc)] += A[IX(i, l, lda)] * t1 + B[IX(i, l, ldb)] * t2;
            }
        }
    } else {
        /* C := alpha*A**T*B + alpha*B**T*A + beta*C, dot products */
        for (int j = 0; j < n; ++j) {
            int i0 = upper ? 0 : j, i1 = upper ? j + 1 : n;
            for (int i = i0; i < i1; ++i) {
                double t1 = 0, t2 = 0;
                for (int l = 0; l < k; ++l) {
                    t1 += A[IX(l, i, lda)] * B[IX(l, j, ldb)];
                    t2 += B[IX(l, i, ldb)] * A[IX(l, j, lda)];
                }
                C[IX(i, j, ldc)] = beta == 0 ? alpha * t1 + alpha * t2
                                             : beta * C[IX(i, j, ldc)] + alpha * t1 + alpha * t2;
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* DTRMM: netlib info (1 side, 2 uplo, 3 transa, 4 diag, 5 m, 6 n,      */
/* 9 lda, 11 ldb); the reference does not build trmm.cc                */
/* (meson.build:87), so the netlib DTRMM definition is the oracle      */
/* ------------------------------------------------------------------ */
int ora_trmm_info(char side, char uplo, char ta, char diag, int m, int n, int lda, int ldb)
{
    return ora_trsm_info(side, uplo, ta, diag, m, n, lda, ldb);
}

void ora_dtrmm(char side, char uplo, char ta, char diag, int m, int n, double alpha,
               const double *A, int lda, double *B, int ldb)
{
    if (ora_trmm_info(side, uplo, ta, diag, m, n, lda, ldb)) return;
    if (m == 0 || n == 0) return;
    int lside = ora_lsame(side, 'L'), upper = ora_lsame(uplo, 'U');
    int notr = ora_lsame(ta, 'N'), nounit = ora_lsame(diag, 'N');
    if (alpha == 0) {
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i) B[IX(i, j, ldb)] = 0;
        return;
    }
    if (lside) {
        if (notr) {
            /* B := alpha*A*B: each column of B, A applied column by column */
            for (int j = 0; j < n; ++j) {
                if (upper) {
                    for (int kk = 0; kk < m; ++kk) {
                        if (B[IX(kk, j, ldb)] == 0) continue;
                        double t = alpha * B[IX(kk, j, ldb)];
                        for (int i = 0; i < kk; ++i) B[IX(i, j, ldb)] += t * A[IX(i, kk, lda)];
                        if (nounit) t *= A[IX(kk, kk, lda)];
                        B[IX(kk, j, ldb)] = t;
                    }
                } else {
                    for (int kk = m - 1; kk >= 0; --kk) {
                        if (B[IX(kk, j, ldb)] == 0) continue;
                        double t = alpha * B[IX(kk, j, ldb)];
                        B[IX(kk, j, ldb)] = t;
                        if (nounit) B[IX(kk, j, ldb)] *= A[IX(kk, kk, lda)];
                        for (int i = kk + 1; i < m; ++i) B[IX(i, j, ldb)] += t * A[IX(i, kk, lda)];
                    }
                }
            }
        } else {
            /* B := alpha*A**T*B: dot products down the columns of A */
            for (int j = 0; j < n; ++j) {
                if (upper) {
                    for (int i = m - 1; i >= 0; --i) {
                        double t = B[IX(i, j, ldb)];
                        if (nounit) t *= A[IX(i, i, lda)];
                        for (int kk = 0; kk < i; ++kk) t += A[IX(kk, i, lda)] * B[IX(kk, j, ldb)];
                        B[IX(i, j, ldb)] = alpha * t;
                    }
                } else {
                    for (int i = 0; i < m; ++i) {
                        double t = B[IX(i, j, ldb)];
                        if (nounit) t *= A[IX(i, i, lda)];
                        for (int kk = i + 1; kk < m; ++kk) t += A[IX(kk, i, lda)] * B[IX(kk, j, ldb)];
                        B[IX(i, j, ldb)] = alpha * t;
                    }
                }
            }
        }
        return;
    }
    if (notr) {
        /* B := alpha*B*A: column j of the result from columns of B */
        if (upper) {
            for (int j = n - 1; j >= 0; --j) {
                double t = alpha;
                if (nounit) t *= A[IX(j, j, lda)];
                for (int i = 0; i < m; ++i) B[IX(i, j, ldb)] *= t;
                for (int kk = 0; kk < j; ++kk) {
                    if (A[IX(kk, j, lda)] == 0) continue;
                    double u = alpha * A[IX(kk, j, lda)];
                    for (int i = 0; i < m; ++i) B[IX(i, j, ldb)] += u * B[IX(i, kk, ldb)];
                }
            }
        } else {
            for (int j = 0; j < n; ++j) {
                double t = alpha;
                if (nounit) t *= A[IX(j, j, lda)];
                for (int i = 0; i < m; ++i) B[IX(i, j, ldb)] *= t;
                for (int kk = j + 1; kk < n; ++kk) {
                    if (A[IX(kk, j, lda)] == 0) continue;
                    double u = alpha * A[IX(kk, j, lda)];
                    for (int i = 0; i < m; ++i) B[IX(i, j, ldb)] += u * B[IX(i, kk, ldb)];
                }
            }
        }
    } else {
        /* B := alpha*B*A**T */
        if (upper) {
            for (int kk = 0; kk < n; ++kk) {
                for (int j = 0; j < kk; ++j) {
                    if (A[IX(j, kk, lda)] == 0) continue;
                    double u = alpha * A[IX(j, kk, lda)];
                    for (int i = 0; i < m; ++i) B[IX(i, j, ldb)] += u * B[IX(i, kk, ldb)];
                }
                double t = alpha;
                if (nounit) t *= A[IX(kk, kk, lda)];
                if (t != 1)
                    for (int i = 0; i < m; ++i) B[IX(i, kk, ldb)] *= t;
            }
        } else {
            for (int kk = n - 1; kk >= 0; --kk) {
                for (int j = kk + 1; j < n; ++j) {
                    if (A[IX(j, kk, lda)] == 0) continue;
                    double u = alpha * A[IX(j, kk, lda)];
                    for (int i = 0; i < m; ++i) B[IX(i, j, ldb)] += u * B[IX(i, kk, ldb)];
                }
                double t = alpha;
                if (nounit) t *= A[IX(kk, kk, lda)];
                if (t != 1)
                    for (int i = 0; i < m; ++i) B[IX(i, kk, ldb)] *= t;
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Level 2                                                             */
/* ------------------------------------------------------------------ */
void ora_dgemv(char trans, int m, int n, double alpha, const double *A, int lda,
               const double *x, int incx, double beta, double *y, int incy)
{
    if (ora_gemv_info(trans, m, n, lda, incx, incy)) return;
    if (m == 0 || n == 0 || (alpha == 0 && beta == 1)) return;
    int notr = ora_lsame(trans, 'N');
    int lenx = notr ? n : m, leny = notr ? m : n;
    int kx = incx > 0 ? 0 : -(lenx - 1) * incx;
    int ky = incy > 0 ? 0 : -(leny - 1) * incy;
    if (beta != 1) {
        int iy = ky;
        for (int i = 0; i < leny; ++i, iy += incy) y[iy] = beta == 0 ? 0.0 : beta * y[iy];
    }
    if (alpha == 0) return;
    if (notr) {
        int jx = kx;
        for (int j = 0; j < n; ++j, jx += incx) {
            double temp = alpha * x[jx];
            int iy = ky;
            for (int i = 0; i < m; ++i, iy += incy) y[iy] += temp * A[IX(i, j, lda)];
        }
    } else {
        int jy = ky;
        for (int j = 0; j < n; ++j, jy += incy) {
            double temp = 0;
            int ix = kx;
            for (int i = 0; i < m; ++i, ix += incx) temp += A[IX(i, j, lda)] * x[ix];
            y[jy] += alpha * temp;
        }
    }
}

void ora_dger(int m, int n, double alpha, const double *x, int incx, const double *y, int incy,
              double *A, int lda)
{
    if (ora_ger_info(m, n, incx, incy, lda)) return;
    if (m == 0 || n == 0 || alpha == 0) return;
    int jy = incy > 0 ? 0 : -(n - 1) * incy;
    int kx = incx > 0 ? 0 : -(m - 1) * incx;
    for (int j = 0; j < n; ++j, jy += incy) {
        if (y[jy] != 0) {
            double temp = alpha * y[jy];
            int ix = kx;
            for (int i = 0; i < m; ++i, ix += incx) A[IX(i, j, lda)] += x[ix] * temp;
        }
    }
}

void ora_dtrsv(char uplo, char trans, char diag, int n, const double *A, int lda, double *x,
               int incx)
{
    if (ora_trsv_info(uplo, trans, diag, n, lda, incx)) return;
    if (n == 0) return;
    int nounit = ora_lsame(diag, 'N'), upper = ora_lsame(uplo, 'U');
    int kx = incx > 0 ? 0 : -(n - 1) * incx;
#define XE(j) x[kx + (j) * incx]
    if (ora_lsame(trans, 'N')) {
        if (upper) {
            for (int j = n - 1; j >= 0; --j)
                if (XE(j) != 0) {
                    if (nounit) XE(j) /= A[IX(j, j, lda)];
                    double temp = XE(j);
                    for (int i = j - 1; i >= 0; --i) XE(i) -= temp * A[IX(i, j, lda)];
                }
        } else {
            for (int j = 0; j < n; ++j)
                if (XE(j) != 0) {
                    if (nounit) XE(j) /= A[IX(j, j, lda)];
                    double temp = XE(j);
                    for (int i = j + 1; i < n; ++i) XE(i) -= temp * A[IX(i, j, lda)];
                }
        }
    } else {
        if (upper) {
            for (int j = 0; j < n; ++j) {
                double temp = XE(j);
                for (int i = 0; i < j; ++i) temp -= A[IX(i, j, lda)] * XE(i);
                if (nounit) temp /= A[IX(j, j, lda)];
                XE(j) = temp;
            }
        } else {
            for (int j = n - 1; j >= 0; --j) {
                double temp = XE(j);
                for (int i = n - 1; i > j; --i) temp -= A[IX(i, j, lda)] * XE(i);
                if (nounit) temp /= A[IX(j, j, lda)];
                XE(j) = temp;
            }
        }
    }
#undef XE
}

/* ------------------------------------------------------------------ */
/* Level 1                                                             */
/* ------------------------------------------------------------------ */
void ora_daxpy(int n, double a, const double *x, int incx, double *y, int incy)
{
    if (n <= 0 || a == 0) return;
    int ix = incx < 0 ? -(n - 1) * incx : 0;
    int iy = incy < 0 ? -(n - 1) * incy : 0;
    for (int i = 0; i < n; ++i, ix += incx, iy += incy) y[iy] += a * x[ix];
}

/* netlib DSCAL/DCOPY/DSWAP/DROT/DASUM (the definitions behind the
 * reference's CBLAS specs blas_level1/scal.cc:42, copy.cc:46, swap.cc:12-29,
 * rot.cc:50, asum.cc:48); element i of a negative-increment vector is at
 * (n-1-i)*|inc| from its start */
static long ora_start(int n, int inc) { return inc < 0 ? (long)(1 - n) * inc : 0; }

void ora_dscal(int n, double a, double *x, int incx)
{
    if (n <= 0 || incx <= 0) return;
    for (int i = 0; i < n; ++i) x[(long)i * incx] = a * x[(long)i * incx];
}

void ora_dcopy(int n, const double *x, int incx, double *y, int incy)
{
    if (n <= 0) return;
    long ix = ora_start(n, incx), iy = ora_start(n, incy);
    for (int i = 0; i < n; ++i, ix += incx, iy += incy) y[iy] = x[ix];
}

void ora_dswap(int n, double *x, int incx, double *y, int incy)
{
    if (n <= 0) return;
    long ix = ora_start(n, incx), iy = ora_start(n, incy);
    for (int i = 0; i < n; ++i, ix += incx, iy += incy) {
        double t = x[ix];
        x[ix] = y[iy];
        y[iy] = t;
    }
}

void ora_drot(int n, double *x, int incx, double *y, int incy, double c, double s)
{
    if (n <= 0) return;
    long ix = ora_start(n, incx), iy = ora_start(n, incy);
    for (int i = 0; i < n; ++i, ix += incx, iy += incy) {
        double t = c * x[ix] + s * y[iy];
        y[iy] = c * y[iy] - s * x[ix];
        x[ix] = t;
    }
}

double ora_dasum(int n, const double *x, int incx)
{
    if (n <= 0 || incx <= 0) return 0.0;
    double t = 0;
    for (int i = 0; i < n; ++i) t += fabs(x[(long)i * incx]);
    return t;
}

double ora_ddot(int n, const double *x, int incx, const double *y, int incy)
{
    double s = 0;
    if (n <= 0) return 0;
    int ix = incx < 0 ? -(n - 1) * incx : 0;
    int iy = incy < 0 ? -(n - 1) * incy : 0;
    for (int i = 0; i < n; ++i, ix += incx, iy += incy) s += x[ix] * y[iy];
    return s;
}

double ora_dnrm2(int n, const double *x, int incx)
{
    /* Scaled sum of squares (the classic netlib DNRM2 recurrence): never
     * overflows or flushes for finite input.  Stride conventions follow the
     * current netlib dnrm2.f90: n <= 0 -> 0, a negative stride walks the
     * vector from its far end (same set of elements), incx == 0 repeats x(1). */
    if (n <= 0) return 0.0;
    if (n == 1) return fabs(x[0]);
    double scale = 0.0, ssq = 1.0;
    long ix = incx < 0 ? -(long)(n - 1) * incx : 0;
    for (int i = 0; i < n; ++i, ix += incx) {
        double v = x[ix];
        if (v != 0) {
            double a = fabs(v);
            if (scale < a) {
                double r = scale / a;
                ssq = 1.0 + ssq * r * r;
                scale = a;
            } else {
                double r = a / scale;
                ssq += r * r;
            }
        }
    }
    return scale * sqrt(ssq);
}

int ora_idamax(int n, const double *x, int incx)
{
    /* netlib IDAMAX: 1-based index of the first maximal |x_i|, strict '>'
     * so ties keep the smaller index and a NaN after position 1 is never
     * selected.  n < 1 or incx <= 0 -> 0. */
    if (n < 1 || incx <= 0) return 0;
    if (n == 1) return 1;
    int best = 1;
    double dmax = fabs(x[0]);
    for (int i = 1; i < n; ++i) {
        double a = fabs(x[(size_t)i * incx]);
        if (a > dmax) {
            best = i + 1;
            dmax = a;
        }
    }
    return best;
}

int ora_idamin(int n, const double *x, int incx)
{
    /* IDAMIN (the extension the reference's cblas_idamin forwards to,
     * blas_level1/amin.cc:40): 1-based index of the first minimal |x_i|, the
     * IDAMAX loop with '<': ties keep the smaller index and a NaN is never
     * selected after position 1.  n < 1 or incx <= 0 -> 0. */
    if (n < 1 || incx <= 0) return 0;
    if (n == 1) return 1;
    int best = 1;
    double dmin = fabs(x[0]);
    for (int i = 1; i < n; ++i) {
        double a = fabs(x[(size_t)i * incx]);
        if (a < dmin) {
            best = i + 1;
            dmin = a;
        }
    }
    return best;
}

void ora_drotm(int n, double *x, int incx, double *y, int incy, const double *param)
{
    /* netlib DROTM (blas_level1/rotm.cc:51 forwards to it): apply the
     * modified Givens H, param = (flag, h11, h21, h12, h22), flag -1: full
     * H, 0: unit diagonal, 1: h12 = 1, h21 = -1, -2: identity. */
    const double flag = param[0];
    if (n <= 0 || flag == -2.0) return;
    const int kx = incx < 0 ? (1 - n) * incx : 0, ky = incy < 0 ? (1 - n) * incy : 0;
    for (int i = 0; i < n; ++i) {
        double *px = x + kx + (long)i * incx, *py = y + ky + (long)i * incy;
        const double w = *px, z = *py;
        if (flag < 0) {
            *px = w * param[1] + z * param[3];
            *py = w * param[2] + z * param[4];
        } else if (flag == 0) {
            *px = w + z * param[3];
            *py = w * param[2] + z;
        } else {
            *px = w * param[1] + z;
            *py = -w + param[4] * z;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Checkers                                                            */
/* ------------------------------------------------------------------ */
double ora_gemm_gauge_ratio(char ta, char tb, int m, int n, int k, double alpha,
                            const double *A, int lda, const double *B, int ldb, double beta,
                            const double *C0, const double *Cgot, int ldc, double eps)
{
    /* tests/netlib/dblat3.f:2496-2546 (DMMCH): expected value and gauge per
     * entry, error ratio = |ct - cc| / (eps * g), 0 where g == 0. */
    int nota = ora_lsame(ta, 'N'), notb = ora_lsame(tb, 'N');
    double worst = 0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            long double ct = 0, g = 0;
            for (int l = 0; l < k; ++l) {
                double a = nota ? A[IX(i, l, lda)] : A[IX(l, i, lda)];
                double b = notb ? B[IX(l, j, ldb)] : B[IX(j, l, ldb)];
                ct += (long double)a * b;
                g += fabsl((long double)a * b);
            }
            ct = (long double)alpha * ct + (long double)beta * C0[IX(i, j, ldc)];
            g = fabsl((long double)alpha) * g + fabsl((long double)beta) * fabs(C0[IX(i, j, ldc)]);
            long double err = fabsl(ct - (long double)Cgot[IX(i, j, ldc)]);
            if (g != 0) {
                double r = (double)(err / ((long double)eps * g));
                if (r > worst) worst = r;
            } else if (err != 0) {
                return INFINITY;
            }
        }
    return worst;
}

uint64_t ora_splitmix64_at(uint64_t seed, uint64_t i)
{
    /* SplitMix64 (Steele, Lea & Flood, OOPSLA 2014): state advances by the
     * golden gamma; element i is the mix of seed + (i+1)*gamma. */
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void ora_fill_uniform(uint64_t seed, int64_t n, double *out)
{
    for (int64_t i = 0; i < n; ++i) {
        uint64_t z = ora_splitmix64_at(seed, (uint64_t)i);
        double u = (double)(z >> 11) * 0x1.0p-53; /* [0,1), exact */
        out[i] = 2.0 * u - 1.0;                   /* [-1,1), exact */
    }
}

/* netlib DSYMV: y := alpha*A*x + beta*y with A symmetric, only the uplo
 * triangle referenced (blas_level2/symv.cc:11-75 forwards here).  Column j
 * adds temp1*A(i,j) to y(i) for the stored off-diagonal i and gathers
 * A(i,j)*x(i) into temp2 for y(j), in netlib's loop order. */
void ora_dsymv(char uplo, int n, double alpha, const double *A, int lda, const double *x,
               int incx, double beta, double *y, int incy)
{
    if (ora_symv_info(uplo, n, lda, incx, incy)) return;
    if (n == 0 || (alpha == 0 && beta == 1)) return;
    int kx = incx > 0 ? 0 : -(n - 1) * incx;
    int ky = incy > 0 ? 0 : -(n - 1) * incy;
    if (beta != 1) {
        int iy = ky;
        for (int i = 0; i < n; ++i, iy += incy) y[iy] = beta == 0 ? 0.0 : beta * y[iy];
    }
    if (alpha == 0) return;
    int jx = kx, jy = ky;
    if (ora_lsame(uplo, 'U')) {
        for (int j = 0; j < n; ++j, jx += incx, jy += incy) {
            double temp1 = alpha * x[jx], temp2 = 0;
            int ix = kx, iy = ky;
            for (int i = 0; i < j; ++i, ix += incx, iy += incy) {
                y[iy] += temp1 * A[IX(i, j, lda)];
                temp2 += A[IX(i, j, lda)] * x[ix];
            }
            y[jy] += temp1 * A[IX(j, j, lda)] + alpha * temp2;
        }
    } else {
        for (int j = 0; j < n; ++j, jx += incx, jy += incy) {
            double temp1 = alpha * x[jx], temp2 = 0;
            y[jy] += temp1 * A[IX(j, j, lda)];
            int ix = jx, iy = jy;
            for (int i = j + 1; i < n; ++i) {
                ix += incx;
                iy += incy;
                y[iy] += temp1 * A[IX(i, j, lda)];
                temp2 += A[IX(i, j, lda)] * x[ix];
            }
            y[jy] += alpha * temp2;
        }
    }
}

/* netlib DSYR: A := alpha*x*x' + A on the uplo triangle (blas_level2/
 * syr.cc:11-68 forwards here); columns with x(j) == 0 are skipped. */
void ora_dsyr(char uplo, int n, double alpha, const double *x, int incx, double *A, int lda)
{
    if (ora_syr_info(uplo, n, incx, lda)) return;
    if (n == 0 || alpha == 0) return;
    int kx = incx > 0 ? 0 : -(n - 1) * incx;
    int upper = ora_lsame(uplo, 'U');
    int jx = kx;
    for (int j = 0; j < n; ++j, jx += incx) {
        if (x[jx] == 0) continue;
        double temp = alpha * x[jx];
        int i0 = upper ? 0 : j, i1 = upper ? j + 1 : n;
        int ix = kx + i0 * incx;
        for (int i = i0; i < i1; ++i, ix += incx) A[IX(i, j, lda)] += x[ix] * temp;
    }
}

/* netlib DSYR2: A := alpha*x*y' + alpha*y*x' + A on the uplo triangle
 * (blas_level2/syr2.cc forwards here); a column with x(j) == y(j) == 0 is
 * skipped, and each element takes A + x(i)*temp1 + y(i)*temp2 in that order. */
void ora_dsyr2(char uplo, int n, double alpha, const double *x, int incx, const double *y,
               int incy, double *A, int lda)
{
    if (ora_syr2_info(uplo, n, incx, incy, lda)) return;
    if (n == 0 || alpha == 0) return;
    const int kx = incx > 0 ? 0 : -(n - 1) * incx, ky = incy > 0 ? 0 : -(n - 1) * incy;
    const int upper = ora_lsame(uplo, 'U');
    for (int j = 0; j < n; ++j) {
        const double xj = x[kx + j * incx], yj = y[ky + j * incy];
        if (xj == 0 && yj == 0) continue;
        const double temp1 = alpha * yj, temp2 = alpha * xj;
        const int i0 = upper ? 0 : j, i1 = upper ? j + 1 : n;
        for (int i = i0; i < i1; ++i)
            A[IX(i, j, lda)] = A[IX(i, j, lda)] + x[kx + i * incx] * temp1 + y[ky + i * incy] * temp2;
    }
}

/* Packed storage (netlib): column j of an upper triangle holds rows 0..j at
 * AP[j(j+1)/2 ..], of a lower one rows j..n-1 at AP[j(2n-j+1)/2 ..]. */
static long pk(int upper, int n, int i, int j)
{
    return upper ? (long)j * (j + 1) / 2 + i : (long)j * (2 * n - j + 1) / 2 + (i - j);
}

/* netlib DSPR: AP := alpha*x*x' + AP (blas_level2/spr.cc forwards here) */
void ora_dspr(char uplo, int n, double alpha, const double *x, int incx, double *AP)
{
    if (ora_spr_info(uplo, n, incx)) return;
    if (n == 0 || alpha == 0) return;
    const int kx = incx > 0 ? 0 : -(n - 1) * incx;
    const int upper = ora_lsame(uplo, 'U');
    for (int j = 0; j < n; ++j) {
        const double xj = x[kx + j * incx];
        if (xj == 0) continue;
        const double temp = alpha * xj;
        const int i0 = upper ? 0 : j, i1 = upper ? j + 1 : n;
        for (int i = i0; i < i1; ++i) AP[pk(upper, n, i, j)] += x[kx + i * incx] * temp;
    }
}

/* netlib DSPR2: AP := alpha*x*y' + alpha*y*x' + AP (blas_level2/spr2.cc) */
void ora_dspr2(char uplo, int n, double alpha, const double *x, int incx, const double *y,
               int incy, double *AP)
{
    if (ora_spr2_info(uplo, n, incx, incy)) return;
    if (n == 0 || alpha == 0) return;
    const int kx = incx > 0 ? 0 : -(n - 1) * incx, ky = incy > 0 ? 0 : -(n - 1) * incy;
    const int upper = ora_lsame(uplo, 'U');
    for (int j = 0; j < n; ++j) {
        const double xj = x[kx + j * incx], yj = y[ky + j * incy];
        if (xj == 0 && yj == 0) continue;
        const double temp1 = alpha * yj, temp2 = alpha * xj;
        const int i0 = upper ? 0 : j, i1 = upper ? j + 1 : n;
        for (int i = i0; i < i1; ++i) {
            const long k = pk(upper, n, i, j);
            AP[k] = AP[k] + x[kx + i * incx] * temp1 + y[ky + i * incy] * temp2;
        }
    }
}

/* Element (i, j) of a triangle in full (lda >= 1, k < 0), netlib band
 * (lda >= k + 1, k >= 0: upper A(k+i-j, j), lower A(i-j, j)) or netlib packed
 * (lda = 0) storage; callers only ask for elements of the stored triangle
 * (within the band).  blo / bhi clip a column's loop to the band. */
struct tri_acc {
    const double *a;
    int lda, upper, n, k;
};
static double tget(const struct tri_acc *t, int i, int j)
{
    if (t->k >= 0) return t->a[IX(t->upper ? t->k + i - j : i - j, j, t->lda)];
    return t->lda ? t->a[IX(i, j, t->lda)] : t->a[pk(t->upper, t->n, i, j)];
}
static int blo(const struct tri_acc *t, int j) { return t->k >= 0 && j - t->k > 0 ? j - t->k : 0; }
static int bhi(const struct tri_acc *t, int j)
{
    return t->k >= 0 && j + t->k < t->n - 1 ? j + t->k : t->n - 1;
}

/* netlib DTRMV / DTPMV: x := op(A)*x, in netlib's loop order (notrans:
 * column sweeps adding x(j)*A(:,j); trans: dot products per element). */
static void trmv_core(const struct tri_acc *t, int trans, int nounit, double *x, int incx)
{
    const int n = t->n, kx = incx > 0 ? 0 : -(n - 1) * incx;
#define XE(j) x[kx + (j) * incx]
    if (!trans) {
        if (t->upper) {
            for (int j = 0; j < n; ++j)
                if (XE(j) != 0) {
                    const double temp = XE(j);
                    for (int i = blo(t, j); i < j; ++i) XE(i) += temp * tget(t, i, j);
                    if (nounit) XE(j) *= tget(t, j, j);
                }
        } else {
            for (int j = n - 1; j >= 0; --j)
                if (XE(j) != 0) {
                    const double temp = XE(j);
                    for (int i = bhi(t, j); i > j; --i) XE(i) += temp * tget(t, i, j);
                    if (nounit) XE(j) *= tget(t, j, j);
                }
        }
    } else {
        if (t->upper) {
            for (int j = n - 1; j >= 0; --j) {
                double temp = XE(j);
                if (nounit) temp *= tget(t, j, j);
                for (int i = j - 1; i >= blo(t, j); --i) temp += tget(t, i, j) * XE(i);
                XE(j) = temp;
            }
        } else {
            for (int j = 0; j < n; ++j) {
                double temp = XE(j);
                if (nounit) temp *= tget(t, j, j);
                for (int i = j + 1; i <= bhi(t, j); ++i) temp += tget(t, i, j) * XE(i);
                XE(j) = temp;
            }
        }
    }
#undef XE
}

void ora_dtrmv(char uplo, char trans, char diag, int n, const double *A, int lda, double *x,
               int incx)
{
    if (ora_trmv_info(uplo, trans, diag, n, lda, incx)) return;
    if (n == 0) return;
    const struct tri_acc t = {A, lda, ora_lsame(uplo, 'U'), n, -1};
    trmv_core(&t, !ora_lsame(trans, 'N'), ora_lsame(diag, 'N'), x, incx);
}

void ora_dtpmv(char uplo, char trans, char diag, int n, const double *AP, double *x, int incx)
{
    if (ora_tpmv_info(uplo, trans, diag, n, incx)) return;
    if (n == 0) return;
    const struct tri_acc t = {AP, 0, ora_lsame(uplo, 'U'), n, -1};
    trmv_core(&t, !ora_lsame(trans, 'N'), ora_lsame(diag, 'N'), x, incx);
}

/* netlib DTPSV / DTBSV: op(A)*x = b in place (column sweeps for notrans, dot
 * products for trans), loops clipped to the band for band storage. */
static void trsv_core(const struct tri_acc *t, int trans, int nounit, double *x, int incx)
{
    const int n = t->n, kx = incx > 0 ? 0 : -(n - 1) * incx;
#define XE(j) x[kx + (j) * incx]
    if (!trans) {
        if (t->upper) {
            for (int j = n - 1; j >= 0; --j)
                if (XE(j) != 0) {
                    if (nounit) XE(j) /= tget(t, j, j);
                    const double temp = XE(j);
                    for (int i = j - 1; i >= blo(t, j); --i) XE(i) -= temp * tget(t, i, j);
                }
        } else {
            for (int j = 0; j < n; ++j)
                if (XE(j) != 0) {
                    if (nounit) XE(j) /= tget(t, j, j);
                    const double temp = XE(j);
                    for (int i = j + 1; i <= bhi(t, j); ++i) XE(i) -= temp * tget(t, i, j);
                }
        }
    } else {
        if (t->upper) {
            for (int j = 0; j < n; ++j) {
                double temp = XE(j);
                for (int i = blo(t, j); i < j; ++i) temp -= tget(t, i, j) * XE(i);
                if (nounit) temp /= tget(t, j, j);
                XE(j) = temp;
            }
        } else {
            for (int j = n - 1; j >= 0; --j) {
                double temp = XE(j);
                for (int i = bhi(t, j); i > j; --i) temp -= tget(t, i, j) * XE(i);
                if (nounit) temp /= tget(t, j, j);
                XE(j) = temp;
            }
        }
    }
#undef XE
}

/* netlib DTPSV: op(A)*x = b in place, packed A (blas_level2/tpsv.cc) */
void ora_dtpsv(char uplo, char trans, char diag, int n, const double *AP, double *x, int incx)
{
    if (ora_tpmv_info(uplo, trans, diag, n, incx)) return;
    if (n == 0) return;
    const struct tri_acc t = {AP, 0, ora_lsame(uplo, 'U'), n, -1};
    trsv_core(&t, !ora_lsame(trans, 'N'), ora_lsame(diag, 'N'), x, incx);
}

/* INFO of netlib DTBMV / DTBSV: 1 uplo, 2 trans, 3 diag, 4 n, 5 k, 7 lda, 9 incx */
int ora_tbmv_info(char uplo, char trans, char diag, int n, int k, int lda, int incx)
{
    if (!ora_lsame(uplo, 'U') && !ora_lsame(uplo, 'L')) return 1;
    if (!ora_lsame(trans, 'N') && !ora_lsame(trans, 'T') && !ora_lsame(trans, 'C')) return 2;
    if (!ora_lsame(diag, 'U') && !ora_lsame(diag, 'N')) return 3;
    if (n < 0) return 4;
    if (k < 0) return 5;
    if (lda < k + 1) return 7;
    if (incx == 0) return 9;
    return 0;
}

/* netlib DTBMV: x := op(A)*x, triangular band A (blas_level2/tbmv.cc) */
void ora_dtbmv(char uplo, char trans, char diag, int n, int k, const double *A, int lda,
               double *x, int incx)
{
    if (ora_tbmv_info(uplo, trans, diag, n, k, lda, incx)) return;
    if (n == 0) return;
    const struct tri_acc t = {A, lda, ora_lsame(uplo, 'U'), n, k};
    trmv_core(&t, !ora_lsame(trans, 'N'), ora_lsame(diag, 'N'), x, incx);
}

/* netlib DTBSV: op(A)*x = b in place, triangular band A (blas_level2/tbsv.cc) */
void ora_dtbsv(char uplo, char trans, char diag, int n, int k, const double *A, int lda,
               double *x, int incx)
{
    if (ora_tbmv_info(uplo, trans, diag, n, k, lda, incx)) return;
    if (n == 0) return;
    const struct tri_acc t = {A, lda, ora_lsame(uplo, 'U'), n, k};
    trsv_core(&t, !ora_lsame(trans, 'N'), ora_lsame(diag, 'N'), x, incx);
}

/* INFO of netlib DGBMV: 1 trans, 2 m, 3 n, 4 kl, 5 ku, 8 lda, 10 incx, 13 incy */
int ora_gbmv_info(char trans, int m, int n, int kl, int ku, int lda, int incx, int incy)
{
    if (!ora_lsame(trans, 'N') && !ora_lsame(trans, 'T') && !ora_lsame(trans, 'C')) return 1;
    if (m < 0) return 2;
    if (n < 0) return 3;
    if (kl < 0) return 4;
    if (ku < 0) return 5;
    if (lda < kl + ku + 1) return 8;
    if (incx == 0) return 10;
    if (incy == 0) return 13;
    return 0;
}

/* netlib DGBMV: y := alpha*op(A)*x + beta*y, general band A stored as
 * A(ku+i-j, j) (blas_level2/gbmv.cc) */
void ora_dgbmv(char trans, int m, int n, int kl, int ku, double alpha, const double *A, int lda,
               const double *x, int incx, double beta, double *y, int incy)
{
    if (ora_gbmv_info(trans, m, n, kl, ku, lda, incx, incy)) return;
    if (m == 0 || n == 0 || (alpha == 0 && beta == 1)) return;
    const int notr = ora_lsame(trans, 'N');
    const int lenx = notr ? n : m, leny = notr ? m : n;
    const int kx = incx > 0 ? 0 : -(lenx - 1) * incx, ky = incy > 0 ? 0 : -(leny - 1) * incy;
    if (beta != 1)
        for (int i = 0; i < leny; ++i) y[ky + i * incy] = beta == 0 ? 0.0 : beta * y[ky + i * incy];
    if (alpha == 0) return;
    for (int j = 0; j < n; ++j) {
        const int i0 = j - ku > 0 ? j - ku : 0, i1 = j + kl < m - 1 ? j + kl : m - 1;
        if (notr) {
            const double temp = alpha * x[kx + j * incx];
            for (int i = i0; i <= i1; ++i) y[ky + i * incy] += temp * A[IX(ku + i - j, j, lda)];
        } else {
            double temp = 0;
            for (int i = i0; i <= i1; ++i) temp += A[IX(ku + i - j, j, lda)] * x[kx + i * incx];
            y[ky + j * incy] += alpha * temp;
        }
    }
}

/* INFO of netlib DSBMV: 1 uplo, 2 n, 3 k, 6 lda, 8 incx, 11 incy */
int ora_sbmv_info(char uplo, int n, int k, int lda, int incx, int incy)
{
    if (!ora_lsame(uplo, 'U') && !ora_lsame(uplo, 'L')) return 1;
    if (n < 0) return 2;
    if (k < 0) return 3;
    if (lda < k + 1) return 6;
    if (incx == 0) return 8;
    if (incy == 0) return 11;
    return 0;
}

/* netlib DSBMV: y := alpha*A*x + beta*y, symmetric band A (blas_level2/sbmv.cc) */
void ora_dsbmv(char uplo, int n, int k, double alpha, const double *A, int lda, const double *x,
               int incx, double beta, double *y, int incy)
{
    if (ora_sbmv_info(uplo, n, k, lda, incx, incy)) return;
    if (n == 0 || (alpha == 0 && beta == 1)) return;
    const int kx = incx > 0 ? 0 : -(n - 1) * incx, ky = incy > 0 ? 0 : -(n - 1) * incy;
    if (beta != 1)
        for (int i = 0; i < n; ++i) y[ky + i * incy] = beta == 0 ? 0.0 : beta * y[ky + i * incy];
    if (alpha == 0) return;
    const struct tri_acc t = {A, lda, ora_lsame(uplo, 'U'), n, k};
    for (int j = 0; j < n; ++j) {
        const double temp1 = alpha * x[kx + j * incx];
        double temp2 = 0;
        if (t.upper) {
            for (int i = blo(&t, j); i < j; ++i) {
                const double a = tget(&t, i, j);
                y[ky + i * incy] += temp1 * a;
                temp2 += a * x[kx + i * incx];
            }
            y[ky + j * incy] += temp1 * tget(&t, j, j) + alpha * temp2;
        } else {
            y[ky + j * incy] += temp1 * tget(&t, j, j);
            for (int i = j + 1; i <= bhi(&t, j); ++i) {
                const double a = tget(&t, i, j);
                y[ky + i * incy] += temp1 * a;
                temp2 += a * x[kx + i * incx];
            }
            y[ky + j * incy] += alpha * temp2;
        }
    }
}

/* netlib DSPMV: y := alpha*A*x + beta*y, packed symmetric A (blas_level2/pmv.cc) */
void ora_dspmv(char uplo, int n, double alpha, const double *AP, const double *x, int incx,
               double beta, double *y, int incy)
{
    if (ora_spmv_info(uplo, n, incx, incy)) return;
    if (n == 0 || (alpha == 0 && beta == 1)) return;
    const int kx = incx > 0 ? 0 : -(n - 1) * incx, ky = incy > 0 ? 0 : -(n - 1) * incy;
    if (beta != 1)
        for (int i = 0; i < n; ++i) y[ky + i * incy] = beta == 0 ? 0.0 : beta * y[ky + i * incy];
    if (alpha == 0) return;
    const int upper = ora_lsame(uplo, 'U');
    for (int j = 0; j < n; ++j) {
        const double temp1 = alpha * x[kx + j * incx];
        double temp2 = 0;
        if (upper) {
            for (int i = 0; i < j; ++i) {
                const double a = AP[pk(1, n, i, j)];
                y[ky + i * incy] += temp1 * a;
                temp2 += a * x[kx + i * incx];
            }
            y[ky + j * incy] += temp1 * AP[pk(1, n, j, j)] + alpha * temp2;
        } else {
            y[ky + j * incy] += temp1 * AP[pk(0, n, j, j)];
            for (int i = j + 1; i < n; ++i) {
                const double a = AP[pk(0, n, i, j)];
                y[ky + i * incy] += temp1 * a;
                temp2 += a * x[kx + i * incx];
            }
            y[ky + j * incy] += alpha * temp2;
        }
    }
}
