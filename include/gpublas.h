/*
 * gpublas.h -- the drop-in surface of libgpublas (MI355X / gfx950).
 *
 * Load it ahead of the host BLAS (LD_PRELOAD=libgpublas_preload.so, or link
 * libgpublas.so before -lblas).  It exports:
 *
 *  1. Fortran-ABI BLAS entry points (all arguments by reference, 32-bit
 *     integers, hidden Fortran string lengths ignored).  Each replaces the
 *     reference's interposed symbol or its CBLAS-only spec:
 *       dgemm_/sgemm_  blas_level3/gemm.cc:143-179 (prototype blas.h:202-215)
 *       dsyrk_/ssyrk_  blas_level3/syrk.cc:148-180 (blas.h:268-274)
 *       dtrsm_/strsm_  blas_level3/trsm.cc:134-172 (blas.h:303-314)
 *       dsymm_/ssymm_  blas_level3/symm.cc:124-156 (blas.h:254-266)
 *       dsyr2k_/ssyr2k_ blas_level3/syr2k.cc:146-180 (blas.h:276-286)
 *       dtrmm_/strmm_  blas_level3/trmm.cc:139-175 (blas.h:290-299; not
 *                      built by the reference, meson.build:87: netlib DTRMM)
 *       dgemv_         blas_level2/gemv.cc:90-98 (blas.h:85-92)
 *       dger_          blas_level2/ger.cc:86-94
 *       dtrsv_         blas_level2/trsv.cc:81-88
 *       dsymv_         blas_level2/symv.cc:11-75 (CBLAS spec; netlib DSYMV)
 *       dsyr_          blas_level2/syr.cc:11-68 (CBLAS spec; netlib DSYR)
 *       dsyr2_ dspr_ dspr2_  blas_level2/syr2.cc, spr.cc, spr2.cc (CBLAS specs;
 *                      netlib DSYR2, DSPR, DSPR2; packed = netlib column packing)
 *       dtrmv_ dtpmv_ dtpsv_ dspmv_  blas_level2/trmv.cc, tpmv.cc, tpsv.cc, pmv.cc
 *                      (CBLAS specs; netlib DTRMV, DTPMV, DTPSV, DSPMV)
 *       dgbmv_ dsbmv_ dtbmv_ dtbsv_  blas_level2/gbmv.cc, sbmv.cc, tbmv.cc, tbsv.cc
 *                      (CBLAS specs; netlib DGBMV, DSBMV, DTBMV, DTBSV; band storage)
 *       daxpy_         blas_level1/axpy.cc:48-50 (blas.h:20-31)
 *       ddot_          blas_level1/dot.cc:44-48 (returns double by value,
 *                      not the reference's void macro blas.h:45-51)
 *       dnrm2_         blas_level1/nrm2.cc:37-41 (returns double)
 *       idamax_        blas_level1/amax.cc:38-43 (returns the 1-based index)
 *       idamin_ drotm_ blas_level1/amin.cc:37-42, rotm.cc:51-56 (IDAMIN: the extension
 *                      cublasIdamin implements; DROTM: netlib)
 *       dscal_ dcopy_ dswap_ drot_ dasum_  netlib definitions behind the
 *                      reference's CBLAS specs blas_level1/scal.cc:42,
 *                      copy.cc:46, swap.cc:12-29, rot.cc:50, asum.cc:48 (dasum_
 *                      returns double)
 *     Illegal arguments call xerbla_("DGEMM ", &info, 6) of the process
 *     (runtime-blas.c:34-52) with netlib INFO codes and return.
 *
 *  2. CBLAS twins cblas_dgemm ... cblas_idamax (cblas.h:57-822), row-major
 *     handled by operand/flag swaps; cblas_idamax returns the 0-based index
 *     (the reference returns cuBLAS's 1-based value, amax.cc:38-42).
 *
 *  3. malloc/calloc/realloc/free (libgpublas_preload.so only): allocations
 *     selected by GPUBLAS_OPTIONS (heuristic=...) become tracked objects whose
 *     contents stay resident in HBM across BLAS calls; replaces the
 *     reference's interposer lib/obj_tracker.c:789-982.
 *
 *  4. The gpublas_* extension API below (options, statistics, explicit
 *     tracked allocation, call batching).
 */
#ifndef GPUBLAS_H
#define GPUBLAS_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- Fortran ABI ---------------- */
void dgemm_(const char *transa, const char *transb, const int *m, const int *n, const int *k,
            const double *alpha, const double *a, const int *lda, const double *b, const int *ldb,
            const double *beta, double *c, const int *ldc);
void sgemm_(const char *transa, const char *transb, const int *m, const int *n, const int *k,
            const float *alpha, const float *a, const int *lda, const float *b, const int *ldb,
            const float *beta, float *c, const int *ldc);
/* complex GEMM (gemm.cc:181-217, F77_gemm(z) / F77_gemm(c)): a, b, c, alpha
 * and beta point to (re, im) pairs, leading dimensions count complex
 * elements; trans 'C' is the conjugate transpose */
void zgemm_(const char *transa, const char *transb, const int *m, const int *n, const int *k,
            const double *alpha, const double *a, const int *lda, const double *b, const int *ldb,
            const double *beta, double *c, const int *ldc);
void cgemm_(const char *transa, const char *transb, const int *m, const int *n, const int *k,
            const float *alpha, const float *a, const int *lda, const float *b, const int *ldb,
            const float *beta, float *c, const int *ldc);
/* the rest of the complex Level-3 (syrk.cc, herk.cc, syr2k.cc, her2k.cc, symm.cc,
 * hemm.cc, trsm.cc: F77_*(c / z); ztrmm_ / ctrmm_ as netlib): complex scalars by
 * address as (re, im) pairs; HERK's alpha and beta and HER2K's beta are real */
void zsyrk_(const char *uplo, const char *trans, const int *n, const int *k, const double *alpha,
            const double *a, const int *lda, const double *beta, double *c, const int *ldc);
void zherk_(const char *uplo, const char *trans, const int *n, const int *k, const double *alpha,
            const double *a, const int *lda, const double *beta, double *c, const int *ldc);
void zsyr2k_(const char *uplo, const char *trans, const int *n, const int *k, const double *alpha,
             const double *a, const int *lda, const double *b, const int *ldb, const double *beta, double *c,
             const int *ldc);
void zher2k_(const char *uplo, const char *trans, const int *n, const int *k, const double *alpha,
             const double *a, const int *lda, const double *b, const int *ldb, const double *beta, double *c,
             const int *ldc);
void zsymm_(const char *side, const char *uplo, const int *m, const int *n, const double *alpha,
            const double *a, const int *lda, const double *b, const int *ldb, const double *beta, double *c,
            const int *ldc);
void zhemm_(const char *side, const char *uplo, const int *m, const int *n, const double *alpha,
            const double *a, const int *lda, const double *b, const int *ldb, const double *beta, double *c,
            const int *ldc);
void ztrmm_(const char *side, const char *uplo, const char *transa, const char *diag, const int *m,
            const int *n, const double *alpha, const double *a, const int *lda, double *b, const int *ldb);
void ztrsm_(const char *side, const char *uplo, const char *transa, const char *diag, const int *m,
            const int *n, const double *alpha, const double *a, const int *lda, double *b, const int *ldb);
void csyrk_(const char *uplo, const char *trans, const int *n, const int *k, const float *alpha,
            const float *a, const int *lda, const float *beta, float *c, const int *ldc);
void cherk_(const char *uplo, const char *trans, const int *n, const int *k, const float *alpha,
            const float *a, const int *lda, const float *beta, float *c, const int *ldc);
void csyr2k_(const char *uplo, const char *trans, const int *n, const int *k, const float *alpha,
             const float *a, const int *lda, const float *b, const int *ldb, const float *beta, float *c,
             const int *ldc);
void cher2k_(const char *uplo, const char *trans, const int *n, const int *k, const float *alpha,
             const float *a, const int *lda, const float *b, const int *ldb, const float *beta, float *c,
             const int *ldc);
void csymm_(const char *side, const char *uplo, const int *m, const int *n, const float *alpha,
            const float *a, const int *lda, const float *b, const int *ldb, const float *beta, float *c,
            const int *ldc);
void chemm_(const char *side, const char *uplo, const int *m, const int *n, const float *alpha,
            const float *a, const int *lda, const float *b, const int *ldb, const float *beta, float *c,
            const int *ldc);
void ctrmm_(const char *side, const char *uplo, const char *transa, const char *diag, const int *m,
            const int *n, const float *alpha, const float *a, const int *lda, float *b, const int *ldb);
void ctrsm_(const char *side, const char *uplo, const char *transa, const char *diag, const int *m,
            const int *n, const float *alpha, const float *a, const int *lda, float *b, const int *ldb);
void dsyrk_(const char *uplo, const char *trans, const int *n, const int *k, const double *alpha,
            const double *a, const int *lda, const double *beta, double *c, const int *ldc);
void ssyrk_(const char *uplo, const char *trans, const int *n, const int *k, const float *alpha,
            const float *a, const int *lda, const float *beta, float *c, const int *ldc);
void dtrsm_(const char *side, const char *uplo, const char *transa, const char *diag, const int *m,
            const int *n, const double *alpha, const double *a, const int *lda, double *b,
            const int *ldb);
void strsm_(const char *side, const char *uplo, const char *transa, const char *diag, const int *m,
            const int *n, const float *alpha, const float *a, const int *lda, float *b,
            const int *ldb);
void dsymm_(const char *side, const char *uplo, const int *m, const int *n, const double *alpha,
            const double *a, const int *lda, const double *b, const int *ldb, const double *beta,
            double *c, const int *ldc);
void ssymm_(const char *side, const char *uplo, const int *m, const int *n, const float *alpha,
            const float *a, const int *lda, const float *b, const int *ldb, const float *beta,
            float *c, const int *ldc);
void dsyr2k_(const char *uplo, const char *trans, const int *n, const int *k, const double *alpha,
             const double *a, const int *lda, const double *b, const int *ldb, const double *beta,
             double *c, const int *ldc);
void ssyr2k_(const char *uplo, const char *trans, const int *n, const int *k, const float *alpha,
             const float *a, const int *lda, const float *b, const int *ldb, const float *beta,
             float *c, const int *ldc);
void dtrmm_(const char *side, const char *uplo, const char *transa, const char *diag, const int *m,
            const int *n, const double *alpha, const double *a, const int *lda, double *b,
            const int *ldb);
void strmm_(const char *side, const char *uplo, const char *transa, const char *diag, const int *m,
            const int *n, const float *alpha, const float *a, const int *lda, float *b,
            const int *ldb);
void dgemv_(const char *trans, const int *m, const int *n, const double *alpha, const double *a,
            const int *lda, const double *x, const int *incx, const double *beta, double *y,
            const int *incy);
void dger_(const int *m, const int *n, const double *alpha, const double *x, const int *incx,
           const double *y, const int *incy, double *a, const int *lda);
void dtrsv_(const char *uplo, const char *trans, const char *diag, const int *n, const double *a,
            const int *lda, double *x, const int *incx);
void dsymv_(const char *uplo, const int *n, const double *alpha, const double *a, const int *lda,
            const double *x, const int *incx, const double *beta, double *y, const int *incy);
void dsyr_(const char *uplo, const int *n, const double *alpha, const double *x, const int *incx,
           double *a, const int *lda);
void dsyr2_(const char *uplo, const int *n, const double *alpha, const double *x,
            const int *incx, const double *y, const int *incy, double *a, const int *lda);
void dspr_(const char *uplo, const int *n, const double *alpha, const double *x, const int *incx,
           double *ap);
void dspr2_(const char *uplo, const int *n, const double *alpha, const double *x,
            const int *incx, const double *y, const int *incy, double *ap);
void dtrmv_(const char *uplo, const char *trans, const char *diag, const int *n, const double *a,
            const int *lda, double *x, const int *incx);
void dtpmv_(const char *uplo, const char *trans, const char *diag, const int *n, const double *ap,
            double *x, const int *incx);
void dtpsv_(const char *uplo, const char *trans, const char *diag, const int *n, const double *ap,
            double *x, const int *incx);
void dspmv_(const char *uplo, const int *n, const double *alpha, const double *ap, const double *x,
            const int *incx, const double *beta, double *y, const int *incy);
void dgbmv_(const char *trans, const int *m, const int *n, const int *kl, const int *ku,
            const double *alpha, const double *a, const int *lda, const double *x, const int *incx,
            const double *beta, double *y, const int *incy);
void dsbmv_(const char *uplo, const int *n, const int *k, const double *alpha, const double *a,
            const int *lda, const double *x, const int *incx, const double *beta, double *y,
            const int *incy);
void dtbmv_(const char *uplo, const char *trans, const char *diag, const int *n, const int *k,
            const double *a, const int *lda, double *x, const int *incx);
void dtbsv_(const char *uplo, const char *trans, const char *diag, const int *n, const int *k,
            const double *a, const int *lda, double *x, const int *incx);
void daxpy_(const int *n, const double *alpha, const double *x, const int *incx, double *y,
            const int *incy);
double ddot_(const int *n, const double *x, const int *incx, const double *y, const int *incy);
double dnrm2_(const int *n, const double *x, const int *incx);
int idamax_(const int *n, const double *x, const int *incx);
int idamin_(const int *n, const double *x, const int *incx);
void drotm_(const int *n, double *x, const int *incx, double *y, const int *incy,
            const double *param);
void dscal_(const int *n, const double *alpha, double *x, const int *incx);
void dcopy_(const int *n, const double *x, const int *incx, double *y, const int *incy);
void dswap_(const int *n, double *x, const int *incx, double *y, const int *incy);
void drot_(const int *n, double *x, const int *incx, double *y, const int *incy, const double *c,
           const double *s);
double dasum_(const int *n, const double *x, const int *incx);
int lsame_(const char *ca, const char *cb);

/* ---------------- CBLAS (cblas.h:21-25 enum values) ---------------- */
typedef enum { CblasRowMajor = 101, CblasColMajor = 102 } CBLAS_LAYOUT;
typedef enum { CblasNoTrans = 111, CblasTrans = 112, CblasConjTrans = 113 } CBLAS_TRANSPOSE;
typedef enum { CblasUpper = 121, CblasLower = 122 } CBLAS_UPLO;
typedef enum { CblasNonUnit = 131, CblasUnit = 132 } CBLAS_DIAG;
typedef enum { CblasLeft = 141, CblasRight = 142 } CBLAS_SIDE;

void cblas_dgemm(CBLAS_LAYOUT layout, CBLAS_TRANSPOSE ta, CBLAS_TRANSPOSE tb, int m, int n, int k,
                 double alpha, const double *a, int lda, const double *b, int ldb, double beta,
                 double *c, int ldc);
void cblas_sgemm(CBLAS_LAYOUT layout, CBLAS_TRANSPOSE ta, CBLAS_TRANSPOSE tb, int m, int n, int k,
                 float alpha, const float *a, int lda, const float *b, int ldb, float beta,
                 float *c, int ldc);
/* complex twins (cblas.h cblas_zgemm / cblas_cgemm): alpha, beta by address */
void cblas_zgemm(CBLAS_LAYOUT layout, CBLAS_TRANSPOSE ta, CBLAS_TRANSPOSE tb, int m, int n, int k,
                 const void *alpha, const void *a, int lda, const void *b, int ldb,
                 const void *beta, void *c, int ldc);
void cblas_cgemm(CBLAS_LAYOUT layout, CBLAS_TRANSPOSE ta, CBLAS_TRANSPOSE tb, int m, int n, int k,
                 const void *alpha, const void *a, int lda, const void *b, int ldb,
                 const void *beta, void *c, int ldc);
void cblas_zsyrk(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, int n, int k,
                 const void *alpha, const void *a, int lda, const void *beta, void *c, int ldc);
void cblas_zherk(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, int n, int k,
                 double alpha, const void *a, int lda, double beta, void *c, int ldc);
void cblas_zsyr2k(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, int n, int k,
                  const void *alpha, const void *a, int lda, const void *b, int ldb,
                  const void *beta, void *c, int ldc);
void cblas_zher2k(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, int n, int k,
                  const void *alpha, const void *a, int lda, const void *b, int ldb, double beta,
                  void *c, int ldc);
void cblas_zsymm(CBLAS_LAYOUT layout, CBLAS_SIDE side, CBLAS_UPLO uplo, int m, int n,
                 const void *alpha, const void *a, int lda, const void *b, int ldb,
                 const void *beta, void *c, int ldc);
void cblas_zhemm(CBLAS_LAYOUT layout, CBLAS_SIDE side, CBLAS_UPLO uplo, int m, int n,
                 const void *alpha, const void *a, int lda, const void *b, int ldb,
                 const void *beta, void *c, int ldc);
void cblas_ztrmm(CBLAS_LAYOUT layout, CBLAS_SIDE side, CBLAS_UPLO uplo, CBLAS_TRANSPOSE ta,
                 CBLAS_DIAG diag, int m, int n, const void *alpha, const void *a, int lda,
                 void *b, int ldb);
void cblas_ztrsm(CBLAS_LAYOUT layout, CBLAS_SIDE side, CBLAS_UPLO uplo, CBLAS_TRANSPOSE ta,
                 CBLAS_DIAG diag, int m, int n, const void *alpha, const void *a, int lda,
                 void *b, int ldb);
void cblas_csyrk(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, int n, int k,
                 const void *alpha, const void *a, int lda, const void *beta, void *c, int ldc);
void cblas_cherk(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, int n, int k,
                 float alpha, const void *a, int lda, float beta, void *c, int ldc);
void cblas_csyr2k(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, int n, int k,
                  const void *alpha, const void *a, int lda, const void *b, int ldb,
                  const void *beta, void *c, int ldc);
void cblas_cher2k(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, int n, int k,
                  const void *alpha, const void *a, int lda, const void *b, int ldb, float beta,
                  void *c, int ldc);
void cblas_csymm(CBLAS_LAYOUT layout, CBLAS_SIDE side, CBLAS_UPLO uplo, int m, int n,
                 const void *alpha, const void *a, int lda, const void *b, int ldb,
                 const void *beta, void *c, int ldc);
void cblas_chemm(CBLAS_LAYOUT layout, CBLAS_SIDE side, CBLAS_UPLO uplo, int m, int n,
                 const void *alpha, const void *a, int lda, const void *b, int ldb,
                 const void *beta, void *c, int ldc);
void cblas_ctrmm(CBLAS_LAYOUT layout, CBLAS_SIDE side, CBLAS_UPLO uplo, CBLAS_TRANSPOSE ta,
                 CBLAS_DIAG diag, int m, int n, const void *alpha, const void *a, int lda,
                 void *b, int ldb);
void cblas_ctrsm(CBLAS_LAYOUT layout, CBLAS_SIDE side, CBLAS_UPLO uplo, CBLAS_TRANSPOSE ta,
                 CBLAS_DIAG diag, int m, int n, const void *alpha, const void *a, int lda,
                 void *b, int ldb);
void cblas_dsyrk(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, int n, int k,
                 double alpha, const double *a, int lda, double beta, double *c, int ldc);
void cblas_dtrsm(CBLAS_LAYOUT layout, CBLAS_SIDE side, CBLAS_UPLO uplo, CBLAS_TRANSPOSE ta,
                 CBLAS_DIAG diag, int m, int n, double alpha, const double *a, int lda, double *b,
                 int ldb);
void cblas_dsymm(CBLAS_LAYOUT layout, CBLAS_SIDE side, CBLAS_UPLO uplo, int m, int n, double alpha,
                 const double *a, int lda, const double *b, int ldb, double beta, double *c,
                 int ldc);
void cblas_dsyr2k(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, int n, int k,
                  double alpha, const double *a, int lda, const double *b, int ldb, double beta,
                  double *c, int ldc);
void cblas_dtrmm(CBLAS_LAYOUT layout, CBLAS_SIDE side, CBLAS_UPLO uplo, CBLAS_TRANSPOSE ta,
                 CBLAS_DIAG diag, int m, int n, double alpha, const double *a, int lda, double *b,
                 int ldb);
/* FP32 Level-3 twins: cblas_ssymm cblas.h:744, cblas_ssyrk :762, cblas_ssyr2k :781,
 * cblas_strmm :800, cblas_strsm :819 (same operand swaps as their FP64 twins) */
void cblas_ssyrk(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, int n, int k,
                 float alpha, const float *a, int lda, float beta, float *c, int ldc);
void cblas_strsm(CBLAS_LAYOUT layout, CBLAS_SIDE side, CBLAS_UPLO uplo, CBLAS_TRANSPOSE ta,
                 CBLAS_DIAG diag, int m, int n, float alpha, const float *a, int lda, float *b,
                 int ldb);
void cblas_ssymm(CBLAS_LAYOUT layout, CBLAS_SIDE side, CBLAS_UPLO uplo, int m, int n, float alpha,
                 const float *a, int lda, const float *b, int ldb, float beta, float *c, int ldc);
void cblas_ssyr2k(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, int n, int k,
                  float alpha, const float *a, int lda, const float *b, int ldb, float beta,
                  float *c, int ldc);
void cblas_strmm(CBLAS_LAYOUT layout, CBLAS_SIDE side, CBLAS_UPLO uplo, CBLAS_TRANSPOSE ta,
                 CBLAS_DIAG diag, int m, int n, float alpha, const float *a, int lda, float *b,
                 int ldb);
void cblas_dgemv(CBLAS_LAYOUT layout, CBLAS_TRANSPOSE trans, int m, int n, double alpha,
                 const double *a, int lda, const double *x, int incx, double beta, double *y,
                 int incy);
void cblas_dger(CBLAS_LAYOUT layout, int m, int n, double alpha, const double *x, int incx,
                const double *y, int incy, double *a, int lda);
void cblas_dtrsv(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, CBLAS_DIAG diag,
                 int n, const double *a, int lda, double *x, int incx);
void cblas_dsymv(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, int n, double alpha, const double *a,
                 int lda, const double *x, int incx, double beta, double *y, int incy);
void cblas_dsyr(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, int n, double alpha, const double *x,
                int incx, double *a, int lda);
void cblas_dsyr2(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, int n, double alpha, const double *x,
                 int incx, const double *y, int incy, double *a, int lda);
void cblas_dspr(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, int n, double alpha, const double *x,
                int incx, double *ap);
void cblas_dspr2(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, int n, double alpha, const double *x,
                 int incx, const double *y, int incy, double *ap);
void cblas_dtrmv(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, CBLAS_DIAG diag,
                 int n, const double *a, int lda, double *x, int incx);
void cblas_dtpmv(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, CBLAS_DIAG diag,
                 int n, const double *ap, double *x, int incx);
void cblas_dtpsv(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, CBLAS_DIAG diag,
                 int n, const double *ap, double *x, int incx);
void cblas_dspmv(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, int n, double alpha, const double *ap,
                 const double *x, int incx, double beta, double *y, int incy);
void cblas_dgbmv(CBLAS_LAYOUT layout, CBLAS_TRANSPOSE trans, int m, int n, int kl, int ku,
                 double alpha, const double *a, int lda, const double *x, int incx, double beta,
                 double *y, int incy);
void cblas_dsbmv(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, int n, int k, double alpha, const double *a,
                 int lda, const double *x, int incx, double beta, double *y, int incy);
void cblas_dtbmv(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, CBLAS_DIAG diag,
                 int n, int k, const double *a, int lda, double *x, int incx);
void cblas_dtbsv(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, CBLAS_DIAG diag,
                 int n, int k, const double *a, int lda, double *x, int incx);
void cblas_daxpy(int n, double alpha, const double *x, int incx, double *y, int incy);
double cblas_ddot(int n, const double *x, int incx, const double *y, int incy);
double cblas_dnrm2(int n, const double *x, int incx);
int cblas_idamax(int n, const double *x, int incx);
int cblas_idamin(int n, const double *x, int incx);
void cblas_drotm(int n, double *x, int incx, double *y, int incy, const double *param);
void cblas_dscal(int n, double alpha, double *x, int incx);
void cblas_dcopy(int n, const double *x, int incx, double *y, int incy);
void cblas_dswap(int n, double *x, int incx, double *y, int incy);
void cblas_drot(int n, double *x, int incx, double *y, int incy, double c, double s);
double cblas_dasum(int n, const double *x, int incx);

/* ---------------- extension API ---------------- */
/* Statistics (the reference keeps hits/misses, runtime-mem.hpp:83,112). */
typedef struct {
    uint64_t calls;          /* BLAS calls entered */
    uint64_t device_calls;   /* calls executed on the GPU */
    uint64_t fallthrough;    /* calls handed to the host BLAS (cutoff) */
    uint64_t hits;           /* operands found resident (tracked) */
    uint64_t misses;         /* operands staged through a temporary device buffer */
    uint64_t h2d_bytes, d2h_bytes;
    uint64_t tracked_objects;/* live tracked objects */
    uint64_t tracked_bytes;
    uint64_t faults;         /* host accesses that pulled a tracked object back */
    uint64_t uploads;        /* tracked objects (re)copied to HBM */
    uint64_t allocs_seen;    /* malloc-family calls counted for ordinals */
    uint64_t async_returns;  /* void calls on resident objects that returned before
                                their kernel finished (host access waits for it) */
    uint64_t evictions;      /* device mirrors dropped to fit the serving budget */
    uint64_t served;         /* calls run by the serving loop */
    uint64_t batches;        /* grouped launches of the serving loop */
    uint64_t max_batch;      /* most calls in one grouped launch */
    uint64_t deferred;       /* admissions postponed until in-flight batches ended */
    uint64_t over_budget;    /* admissions past the budget (nothing left to wait for) */
    uint64_t resident_bytes; /* HBM held by device mirrors now (not reset) */
    uint64_t serve_wait_ns;  /* served calls: summed submission -> launch time */
    uint64_t serve_run_ns;   /* served calls: summed launch -> completion seen */
    uint64_t serve_launch_ns;/* engine time spent forming and launching batches */
    uint64_t resident;       /* reductions served by the resident server (option resident) */
    uint64_t io_syncs;       /* tracked objects brought to the host for a read/write/fread...
                                call on their memory (device-newer download or view opened) */
    uint64_t handler_reinstalls; /* SIGSEGV handler found replaced (not through
                                sigaction/signal) and installed again */
} gpublas_stats_t;
void gpublas_get_stats(gpublas_stats_t *out);
/* Test hooks for the coherence machinery without a device: set the
 * residency state of a tracked allocation that has no device copy yet
 * (0 HOST: view read-write, 1 SHARED: read-only, 2 DEVICE: inaccessible;
 * returns 0, or -1 if p does not start such an allocation), and read it
 * back (-1 if p is not tracked).  No counterpart in the reference, whose
 * managed memory has no such states. */
int gpublas_debug_set_residency(void *p, int state);
int gpublas_debug_residency(const void *p);
void gpublas_reset_stats(void);
/* Wait for every kernel the calling thread has launched (calls on resident
 * operands may return before their kernel ends; see async_returns). */
void gpublas_sync(void);

/* Options string (same syntax as GPUBLAS_OPTIONS); returns 0 if all parsed. */
int gpublas_set_options(const char *opts);
/* GEMM host fall-through cutoff: leading dimensions below it go to the host
 * BLAS (reference rule gemm.cc:129-141, default 512); 0 = always the GPU;
 * a negative value (GPUBLAS_CUTOFF_AUTO) = the measured crossover table. */
void gpublas_set_cutoff(int cutoff);
int gpublas_get_cutoff(void);

/* Measured host/GPU crossover (cutoff=auto, GPUBLAS_CUTOFF_AUTO as the
 * cutoff).  The reference has no counterpart: it replaces the fixed rule of
 * gemm.cc:129-141 where its TODO.md:7 asks for the host BLAS to be picked
 * dynamically.  A call with no resident operand runs on the GPU when its
 * flops reach the routine's min_flops and its flops per operand byte reach
 * min_intensity.  Routine names: dgemm sgemm dsyrk ssyrk dsyr2k ssyr2k dsymm
 * ssymm dtrsm strsm dtrmm strmm dgemv dger dtrsv daxpy (all streaming Level-1
 * calls) ddot (all Level-1 reductions).  Return 0, or -1 for an unknown
 * routine or a negative / NaN threshold. */
#define GPUBLAS_CUTOFF_AUTO (-1)
int gpublas_set_crossover(const char *routine, double min_flops, double min_intensity);
int gpublas_get_crossover(const char *routine, double *min_flops, double *min_intensity);
/* 1 if the table sends a call of this routine with these flops and operand
 * bytes to the GPU, 0 if to the host, -1 for an unknown routine. */
int gpublas_crossover_gpu(const char *routine, double flops, double bytes);
/* Time square problems (n up to max_n for Level 3, 2*max_n for Level 2,
 * 4096*max_n elements for Level 1) on the host BLAS and on the GPU with
 * operands that are not resident, install the resulting table and, if path is
 * not NULL, write it there (the measurements as comments).  Needs the GPU;
 * takes seconds.  Returns 0, -1 (max_n < 16), -2 (out of memory) or -3 (cannot
 * write path). */
int gpublas_tune_crossover(const char *path, int max_n);

/* Explicit tracked allocation: host-visible memory whose contents stay
 * resident in HBM between BLAS calls.  Host reads/writes are coherent (a
 * host touch after a GPU write pulls the object back, a host write after an
 * upload invalidates the device copy). */
void *gpublas_malloc(size_t bytes);
void *gpublas_calloc(size_t nmemb, size_t size);
void gpublas_free(void *p);
int gpublas_is_tracked(const void *p);
/* Device address of a tracked pointer after making the device copy valid
 * (for callers that launch their own kernels); NULL if untracked. */
void *gpublas_device_ptr(const void *p);

/* Number of the last xerbla report on this thread and its routine name. */
int gpublas_last_xerbla(char name[8]);

/* Kernel timing with HIP events (profile option): summed ms, launch count. */
int gpublas_kernel_time(const char *kernel, double *total_ms);

/* 1 when a gfx950 device is usable. */
int gpublas_device_ready(void);

#ifdef __cplusplus
}
#endif
#endif
