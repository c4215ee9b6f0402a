/*
 * oracle/refblas.h -- CPU restatement of the reference BLAS semantics on the
 * interposed hot path.  TEST INFRASTRUCTURE ONLY: tests/, the smoke check and
 * the CPU-baseline leg of bench.py load this as the checker; the product
 * library never links, loads or calls it.
 *
 * The reference (Prince781/libgpublas) contains no arithmetic of its own: the
 * device work is vendor BLAS and the fall-through is the next host BLAS in
 * link order (runtime-blas.c:60-69, gemm.cc:129-141).  What the reference
 * fixes is (1) netlib argument checking and quick returns (gemm.cc:87-127,
 * syrk.cc:78-146, trsm.cc:76-132, symm.cc:81-122) and (2) netlib BLAS
 * arithmetic of the host library it falls through to.  Both are restated
 * here as plain loops in the netlib loop order.
 *
 * All matrices are column-major, 0-based pointers, Fortran leading dims.
 * Every routine takes its scalars by value (this is a checker API, not the
 * Fortran ABI); info functions return the netlib INFO code (0 = legal).
 */
#ifndef GPUBLAS_ORACLE_REFBLAS_H
#define GPUBLAS_ORACLE_REFBLAS_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

int ora_lsame(char a, char b);

/* ---- argument checks (netlib INFO codes) ---- */
int ora_gemm_info(char ta, char tb, int m, int n, int k, int lda, int ldb, int ldc);
int ora_syrk_info(char uplo, char trans, int n, int k, int lda, int ldc);
int ora_trsm_info(char side, char uplo, char ta, char diag, int m, int n, int lda, int ldb);
int ora_symm_info(char side, char uplo, int m, int n, int lda, int ldb, int ldc);
int ora_syr2k_info(char uplo, char trans, int n, int k, int lda, int ldb, int ldc);
int ora_trmm_info(char side, char uplo, char ta, char diag, int m, int n, int lda, int ldb);
void ora_dsyr2k(char uplo, char trans, int n, int k, double alpha, const double *A, int lda,
                const double *B, int ldb, double beta, double *C, int ldc);
void ora_dtrmm(char side, char uplo, char ta, char diag, int m, int n, double alpha,
               const double *A, int lda, double *B, int ldb);
int ora_gemv_info(char trans, int m, int n, int lda, int incx, int incy);
int ora_ger_info(int m, int n, int incx, int incy, int lda);
int ora_trsv_info(char uplo, char trans, char diag, int n, int lda, int incx);
int ora_symv_info(char uplo, int n, int lda, int incx, int incy);
int ora_syr_info(char uplo, int n, int incx, int lda);
int ora_syr2_info(char uplo, int n, int incx, int incy, int lda);
int ora_spr_info(char uplo, int n, int incx);
int ora_spr2_info(char uplo, int n, int incx, int incy);
int ora_trmv_info(char uplo, char trans, char diag, int n, int lda, int incx);
int ora_tpmv_info(char uplo, char trans, char diag, int n, int incx); /* also DTPSV */
int ora_spmv_info(char uplo, int n, int incx, int incy);
int ora_tbmv_info(char uplo, char trans, char diag, int n, int k, int lda, int incx); /* + DTBSV */
int ora_gbmv_info(char trans, int m, int n, int kl, int ku, int lda, int incx, int incy);
int ora_sbmv_info(char uplo, int n, int k, int lda, int incx, int incy);

/* ---- Level 3 ---- */
void ora_dgemm(char ta, char tb, int m, int n, int k, double alpha, const double *A, int lda,
               const double *B, int ldb, double beta, double *C, int ldc);
void ora_sgemm(char ta, char tb, int m, int n, int k, float alpha, const float *A, int lda,
               const float *B, int ldb, float beta, float *C, int ldc);
/* complex GEMM (netlib ZGEMM / CGEMM): operands and scalars are (re, im) pairs */
void ora_zgemm(char ta, char tb, int m, int n, int k, const double *alpha, const double *A, int lda,
               const double *B, int ldb, const double *beta, double *C, int ldc);
void ora_cgemm(char ta, char tb, int m, int n, int k, const float *alpha, const float *A, int lda,
               const float *B, int ldb, const float *beta, float *C, int ldc);
/* the rest of the complex Level-3 (netlib semantics; HERK alpha / beta and HER2K beta real) */
#define ORA_CPLX_DECL(P, R)                                                                   \
    void ora_##P##syrk(char uplo, char trans, int n, int k, const R *alpha, const R *A, int lda, \
                       const R *beta, R *C, int ldc);                                          \
    void ora_##P##herk(char uplo, char trans, int n, int k, const R *alpha, const R *A, int lda, \
                       const R *beta, R *C, int ldc);                                          \
    void ora_##P##syr2k(char uplo, char trans, int n, int k, const R *alpha, const R *A,       \
                        int lda, const R *B, int ldb, const R *beta, R *C, int ldc);           \
    void ora_##P##her2k(char uplo, char trans, int n, int k, const R *alpha, const R *A,       \
                        int lda, const R *B, int ldb, const R *beta, R *C, int ldc);           \
    void ora_##P##symm(char side, char uplo, int m, int n, const R *alpha, const R *A, int lda, \
                       const R *B, int ldb, const R *beta, R *C, int ldc);                     \
    void ora_##P##hemm(char side, char uplo, int m, int n, const R *alpha, const R *A, int lda, \
                       const R *B, int ldb, const R *beta, R *C, int ldc);                     \
    void ora_##P##trmm(char side, char uplo, char transa, char diag, int m, int n,             \
                       const R *alpha, const R *A, int lda, R *B, int ldb);                    \
    void ora_##P##trsm(char side, char uplo, char transa, char diag, int m, int n,             \
                       const R *alpha, const R *A, int lda, R *B, int ldb);
ORA_CPLX_DECL(z, double)
ORA_CPLX_DECL(c, float)
#undef ORA_CPLX_DECL
void ora_dsyrk(char uplo, char trans, int n, int k, double alpha, const double *A, int lda,
               double beta, double *C, int ldc);
void ora_ssyrk(char uplo, char trans, int n, int k, float alpha, const float *A, int lda,
               float beta, float *C, int ldc);
void ora_dtrsm(char side, char uplo, char ta, char diag, int m, int n, double alpha,
               const double *A, int lda, double *B, int ldb);
void ora_strsm(char side, char uplo, char ta, char diag, int m, int n, float alpha,
               const float *A, int lda, float *B, int ldb);
void ora_dsymm(char side, char uplo, int m, int n, double alpha, const double *A, int lda,
               const double *B, int ldb, double beta, double *C, int ldc);
void ora_ssymm(char side, char uplo, int m, int n, float alpha, const float *A, int lda,
               const float *B, int ldb, float beta, float *C, int ldc);

/* ---- Level 2 ---- */
void ora_dgemv(char trans, int m, int n, double alpha, const double *A, int lda,
               const double *x, int incx, double beta, double *y, int incy);
void ora_dger(int m, int n, double alpha, const double *x, int incx, const double *y, int incy,
              double *A, int lda);
void ora_dtrsv(char uplo, char trans, char diag, int n, const double *A, int lda, double *x,
               int incx);

/* ---- Level 1 ---- */
void ora_dscal(int n, double a, double *x, int incx);
void ora_dcopy(int n, const double *x, int incx, double *y, int incy);
void ora_dswap(int n, double *x, int incx, double *y, int incy);
void ora_drot(int n, double *x, int incx, double *y, int incy, double c, double s);
double ora_dasum(int n, const double *x, int incx);
void ora_dsymv(char uplo, int n, double alpha, const double *A, int lda, const double *x,
               int incx, double beta, double *y, int incy);
void ora_dsyr(char uplo, int n, double alpha, const double *x, int incx, double *A, int lda);
void ora_dsyr2(char uplo, int n, double alpha, const double *x, int incx, const double *y,
               int incy, double *A, int lda);
void ora_dspr(char uplo, int n, double alpha, const double *x, int incx, double *AP);
void ora_dtrmv(char uplo, char trans, char diag, int n, const double *A, int lda, double *x,
               int incx);
void ora_dtpmv(char uplo, char trans, char diag, int n, const double *AP, double *x, int incx);
void ora_dtpsv(char uplo, char trans, char diag, int n, const double *AP, double *x, int incx);
void ora_dspmv(char uplo, int n, double alpha, const double *AP, const double *x, int incx,
               double beta, double *y, int incy);
void ora_dtbmv(char uplo, char trans, char diag, int n, int k, const double *A, int lda,
               double *x, int incx);
void ora_dtbsv(char uplo, char trans, char diag, int n, int k, const double *A, int lda,
               double *x, int incx);
void ora_dgbmv(char trans, int m, int n, int kl, int ku, double alpha, const double *A, int lda,
               const double *x, int incx, double beta, double *y, int incy);
void ora_dsbmv(char uplo, int n, int k, double alpha, const double *A, int lda, const double *x,
               int incx, double beta, double *y, int incy);
void ora_dspr2(char uplo, int n, double alpha, const double *x, int incx, const double *y,
               int incy, double *AP);
void ora_daxpy(int n, double a, const double *x, int incx, double *y, int incy);
double ora_ddot(int n, const double *x, int incx, const double *y, int incy);
double ora_dnrm2(int n, const double *x, int incx);
int ora_idamax(int n, const double *x, int incx);
int ora_idamin(int n, const double *x, int incx);
void ora_drotm(int n, double *x, int incx, double *y, int incy, const double *param);

/* ---- checkers ---- */
/* netlib DMMCH gauge ratio for C = alpha*op(A)*op(B) + beta*C0:
 * max_ij |Cgot-Cexp| / (eps * G_ij), G_ij = |alpha| sum_l |op(A)_il||op(B)_lj| + |beta||C0_ij|.
 * Cexp is recomputed here in long double. */
double ora_gemm_gauge_ratio(char ta, char tb, int m, int n, int k, double alpha,
                            const double *A, int lda, const double *B, int ldb, double beta,
                            const double *C0, const double *Cgot, int ldc, double eps);

/* ---- deterministic inputs: SplitMix64 stream, element i = mix(seed + (i+1)*gamma) ---- */
void ora_fill_uniform(uint64_t seed, int64_t n, double *out); /* uniform [-1,1) */
uint64_t ora_splitmix64_at(uint64_t seed, uint64_t i);

#ifdef __cplusplus
}
#endif
#endif
