/*
 * level3.c -- Fortran-ABI Level-3 entry points: ?gemm_, ?syrk_, ?syr2k_, ?trmm_, ?trsm_,
 * ?symm_ for real single and double precision, and the complex Level-3 (z / c: gemm,
 * syrk, herk, syr2k, her2k, symm, hemm, trmm, trsm).
 *
 * Each entry follows the reference's three steps (blas_level3/gemm.cc:87-179
 * and the syrk/trsm/symm twins): netlib argument check with xerbla, quick
 * return, then either the host fall-through (the GEMM cutoff of
 * gemm.cc:129-141; under cutoff=auto the measured crossover of route.c, for
 * every routine) or the device path: stage the operands (tracked objects in place, others
 * through pool buffers), launch the gfx950 kernel, copy results back, and
 * synchronise before returning so results are visible to the caller (the
 * reference skips that sync on some devices, runtime.h:33-34).
 *
 * The alpha == 0 / k == 0 paths follow netlib (C := beta*C, exact zeros for
 * beta == 0), not the reference's C *= beta (gemm.cc:121-126) nor its double
 * scaling in SYRK (syrk.cc:122-145).  They run on the host when C is
 * untracked and on the GPU when C is device-resident.
 */
#include "internal.h"
#include <stddef.h>

#define MAXI(a, b) ((a) > (b) ? (a) : (b))
#define IDX(i, j, ld) ((size_t)(j) * (size_t)(ld) + (size_t)(i))

/* DGEMM / DSYRK device dispatch: plain FP64 MFMA, or the opt-in
 * split-precision scheme (GPUBLAS_OPTIONS split=<slices>). */
static int dgemm_dev(void *s, char ta, char tb, int m, int n, int k, double alpha,
                     const double *A, int lda, const double *B, int ldb, double beta, double *C,
                     int ldc)
{
    if (gb_opt.split > 0)
        return gpublas_dev_dgemm_split_fmt(s, gb_opt.split, gb_opt.split_fmt, ta, tb, m, n, k,
                                           alpha, A, lda, B, ldb, beta, C, ldc);
    return gpublas_dev_dgemm(s, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

static int dsyrk_dev(void *s, char uplo, char trans, int n, int k, double alpha, const double *A,
                     int lda, double beta, double *C, int ldc)
{
    if (gb_opt.split > 0)
        return gpublas_dev_dsyrk_split_fmt(s, gb_opt.split, gb_opt.split_fmt, uplo, trans, n, k,
                                           alpha, A, lda, beta, C, ldc);
    return gpublas_dev_dsyrk(s, uplo, trans, n, k, alpha, A, lda, beta, C, ldc);
}

/* ------------------------------------------------------------------ */
/* GEMM                                                                */
/* ------------------------------------------------------------------ */
#define DEFINE_GEMM(PRE, T, NAME6, DEVFN, SERVEFN, RK)                                                    \
    void PRE##gemm_(const char *transa, const char *transb, const int *m, const int *n,          \
                    const int *k, const T *alpha, const T *a, const int *lda, const T *b,        \
                    const int *ldb, const T *beta, T *c, const int *ldc)                          \
    {                                                                                            \
        gb_enter();                                                                              \
        GB_COUNT(calls, 1);                                                                      \
        const int nota = gb_lsame(*transa, 'N'), notb = gb_lsame(*transb, 'N');                  \
        const int nrowa = nota ? *m : *k, nrowb = notb ? *k : *n;                                \
        int info = 0;                                                                            \
        if (!nota && !gb_lsame(*transa, 'C') && !gb_lsame(*transa, 'T')) info = 1;               \
        else if (!notb && !gb_lsame(*transb, 'C') && !gb_lsame(*transb, 'T')) info = 2;          \
        else if (*m < 0) info = 3;                                                               \
        else if (*n < 0) info = 4;                                                               \
        else if (*k < 0) info = 5;                                                               \
        else if (*lda < MAXI(1, nrowa)) info = 8;                                                \
        else if (*ldb < MAXI(1, nrowb)) info = 10;                                               \
        else if (*ldc < MAXI(1, *m)) info = 13;                                                  \
        if (info) {                                                                              \
            gb_xerbla(NAME6, info);                                                              \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        if (*m == 0 || *n == 0 || ((*alpha == 0 || *k == 0) && *beta == 1)) {                    \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        const int cut = gb_opt.cutoff;                                                           \
        const double es_ = sizeof(T), mm_ = *m, nn_ = *n, kk_ = *k;                              \
        if (gb_host_route(RK, cut > 0 && (*lda < cut || *ldb < cut || *ldc < cut),             \
                          2 * mm_ * nn_ * kk_, es_ * (mm_ * kk_ + kk_ * nn_ + 2 * mm_ * nn_), a,  \
                          b, c)) {                                                               \
            typedef void (*fn_t)(const char *, const char *, const int *, const int *,           \
                                 const int *, const T *, const T *, const int *, const T *,      \
                                 const int *, const T *, T *, const int *);                      \
            fn_t f = (fn_t)gb_host_fn(#PRE "gemm_");                                             \
            GB_COUNT(fallthrough, 1);                                                            \
            f(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);                     \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        const int scale_only = (*alpha == 0 || *k == 0);                                         \
        if (scale_only && !gb_ptr_tracked(c)) {                                                  \
            for (int j = 0; j < *n; ++j)                                                         \
                for (int i = 0; i < *m; ++i)                                                     \
                    c[IDX(i, j, *ldc)] = *beta == 0 ? (T)0 : *beta * c[IDX(i, j, *ldc)];         \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        if (!scale_only && gb_opt.serve &&                                                       \
            SERVEFN(*transa, *transb, *m, *n, *k, *alpha, a, *lda, b, *ldb, *beta, c, *ldc)) {    \
            GB_COUNT(device_calls, 1);                                                           \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        gb_require_device();                                                                     \
        void *s = gpublas_dev_stream();                                                          \
        gb_operand ops[3] = {{0}};                                                               \
        if (!scale_only) {                                                                       \
            gb_acquire_matrix(&ops[0], a, *lda, nrowa, nota ? *k : *m, sizeof(T), GB_IN, s);     \
            gb_acquire_matrix(&ops[1], b, *ldb, nrowb, notb ? *n : *k, sizeof(T), GB_IN, s);     \
        }                                                                                        \
        gb_acquire_matrix(&ops[2], c, *ldc, *m, *n, sizeof(T), *beta == 0 ? GB_OUT : GB_INOUT,  \
                          s);                                                                    \
        gb_check(DEVFN(s, *transa, *transb, *m, *n, *k, *alpha, (const T *)ops[0].dev, *lda,     \
                       (const T *)ops[1].dev, *ldb, *beta, (T *)ops[2].dev, *ldc),               \
                 #PRE "gemm launch");                                                            \
        gb_finish(s, ops, 3);                                                                    \
        GB_COUNT(device_calls, 1);                                                               \
        gb_leave();                                                                              \
    }

DEFINE_GEMM(d, double, "DGEMM ", dgemm_dev, gb_serve_dgemm, GB_RK_DGEMM)
DEFINE_GEMM(s, float, "SGEMM ", gpublas_dev_sgemm, gb_serve_sgemm, GB_RK_SGEMM)

/* ------------------------------------------------------------------ */
/* complex GEMM (zgemm_ / cgemm_, gemm.cc:181-217): operands, alpha and    */
/* beta are (re, im) pairs; 'C' is the conjugate transpose.  Routing uses  */
/* the real GEMM's crossover row with the complex flop and byte counts.    */
/* ------------------------------------------------------------------ */
#define DEFINE_CGEMM(PRE, R, NAME6, DEVFN, RK)                                                   \
    void PRE##gemm_(const char *transa, const char *transb, const int *m, const int *n,          \
                    const int *k, const R *alpha, const R *a, const int *lda, const R *b,        \
                    const int *ldb, const R *beta, R *c, const int *ldc)                          \
    {                                                                                            \
        gb_enter();                                                                              \
        GB_COUNT(calls, 1);                                                                      \
        const int nota = gb_lsame(*transa, 'N'), notb = gb_lsame(*transb, 'N');                  \
        const int nrowa = nota ? *m : *k, nrowb = notb ? *k : *n;                                \
        int info = 0;                                                                            \
        if (!nota && !gb_lsame(*transa, 'C') && !gb_lsame(*transa, 'T')) info = 1;               \
        else if (!notb && !gb_lsame(*transb, 'C') && !gb_lsame(*transb, 'T')) info = 2;          \
        else if (*m < 0) info = 3;                                                               \
        else if (*n < 0) info = 4;                                                               \
        else if (*k < 0) info = 5;                                                               \
        else if (*lda < MAXI(1, nrowa)) info = 8;                                                \
        else if (*ldb < MAXI(1, nrowb)) info = 10;                                               \
        else if (*ldc < MAXI(1, *m)) info = 13;                                                  \
        if (info) {                                                                              \
            gb_xerbla(NAME6, info);                                                              \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        const int a0 = alpha[0] == 0 && alpha[1] == 0, b1 = beta[0] == 1 && beta[1] == 0;        \
        const int b0 = beta[0] == 0 && beta[1] == 0;                                             \
        if (*m == 0 || *n == 0 || ((a0 || *k == 0) && b1)) {                                     \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        const int cut = gb_opt.cutoff;                                                           \
        const double es_ = 2 * sizeof(R), mm_ = *m, nn_ = *n, kk_ = *k;                          \
        if (gb_host_route(RK, cut > 0 && (*lda < cut || *ldb < cut || *ldc < cut),             \
                          8 * mm_ * nn_ * kk_, es_ * (mm_ * kk_ + kk_ * nn_ + 2 * mm_ * nn_), a,  \
                          b, c)) {                                                               \
            typedef void (*fn_t)(const char *, const char *, const int *, const int *,           \
                                 const int *, const R *, const R *, const int *, const R *,      \
                                 const int *, const R *, R *, const int *);                      \
            fn_t f = (fn_t)gb_host_fn(#PRE "gemm_");                                             \
            GB_COUNT(fallthrough, 1);                                                            \
            f(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);                     \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        const int scale_only = (a0 || *k == 0);                                                  \
        if (scale_only && !gb_ptr_tracked(c)) {                                                  \
            for (int j = 0; j < *n; ++j)                                                         \
                for (int i = 0; i < *m; ++i) {                                                   \
                    R *e = c + 2 * IDX(i, j, *ldc);                                              \
                    const R cr = e[0], ci = e[1];                                                \
                    e[0] = b0 ? (R)0 : beta[0] * cr - beta[1] * ci;                              \
                    e[1] = b0 ? (R)0 : beta[0] * ci + beta[1] * cr;                              \
                }                                                                                \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        gb_require_device();                                                                     \
        void *s = gpublas_dev_stream();                                                          \
        gb_operand ops[3] = {{0}};                                                               \
        if (!scale_only) {                                                                       \
            gb_acquire_matrix(&ops[0], a, *lda, nrowa, nota ? *k : *m, 2 * sizeof(R), GB_IN, s); \
            gb_acquire_matrix(&ops[1], b, *ldb, nrowb, notb ? *n : *k, 2 * sizeof(R), GB_IN, s); \
        }                                                                                        \
        gb_acquire_matrix(&ops[2], c, *ldc, *m, *n, 2 * sizeof(R), b0 ? GB_OUT : GB_INOUT, s);   \
        gb_check(DEVFN(s, *transa, *transb, *m, *n, *k, alpha[0], alpha[1],                      \
                       (const R *)ops[0].dev, *lda, (const R *)ops[1].dev, *ldb, beta[0],        \
                       beta[1], (R *)ops[2].dev, *ldc),                                          \
                 #PRE "gemm launch");                                                            \
        gb_finish(s, ops, 3);                                                                    \
        GB_COUNT(device_calls, 1);                                                               \
        gb_leave();                                                                              \
    }

DEFINE_CGEMM(z, double, "ZGEMM ", gpublas_dev_zgemm, GB_RK_DGEMM)
DEFINE_CGEMM(c, float, "CGEMM ", gpublas_dev_cgemm, GB_RK_SGEMM)

/* ------------------------------------------------------------------ */
/* SYRK                                                                */
/* ------------------------------------------------------------------ */
#define DEFINE_SYRK(PRE, T, NAME6, DEVFN, RK)                                                    \
    void PRE##syrk_(const char *uplo, const char *trans, const int *n, const int *k,             \
                    const T *alpha, const T *a, const int *lda, const T *beta, T *c,             \
                    const int *ldc)                                                              \
    {                                                                                            \
        gb_enter();                                                                              \
        GB_COUNT(calls, 1);                                                                      \
        const int notr = gb_lsame(*trans, 'N');                                                  \
        const int nrowa = notr ? *n : *k;                                                        \
        const int upper = gb_lsame(*uplo, 'U');                                                  \
        int info = 0;                                                                            \
        if (!upper && !gb_lsame(*uplo, 'L')) info = 1;                                           \
        else if (!notr && !gb_lsame(*trans, 'T') && !gb_lsame(*trans, 'C')) info = 2;            \
        else if (*n < 0) info = 3;                                                               \
        else if (*k < 0) info = 4;                                                               \
        else if (*lda < MAXI(1, nrowa)) info = 7;                                                \
        else if (*ldc < MAXI(1, *n)) info = 10;                                                  \
        if (info) {                                                                              \
            gb_xerbla(NAME6, info);                                                              \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        if (*n == 0 || ((*alpha == 0 || *k == 0) && *beta == 1)) {                               \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        {                                                                                        \
            const double es_ = sizeof(T), nn_ = *n, kk_ = *k;                                    \
            if (gb_host_route(RK, 0, nn_ * nn_ * kk_, es_ * (nn_ * kk_ + 2 * nn_ * nn_), a, c,   \
                              NULL)) {                                                           \
                typedef void (*fn_t)(const char *, const char *, const int *, const int *,       \
                                     const T *, const T *, const int *, const T *, T *,          \
                                     const int *);                                               \
                GB_COUNT(fallthrough, 1);                                                        \
                ((fn_t)gb_host_fn(#PRE "syrk_"))(uplo, trans, n, k, alpha, a, lda, beta, c, ldc); \
                gb_leave();                                                                      \
                return;                                                                          \
            }                                                                                    \
        }                                                                                        \
        const int scale_only = (*alpha == 0 || *k == 0);                                         \
        if (scale_only && !gb_ptr_tracked(c)) {                                                  \
            for (int j = 0; j < *n; ++j) {                                                       \
                const int i0 = upper ? 0 : j, i1 = upper ? j + 1 : *n;                           \
                for (int i = i0; i < i1; ++i)                                                    \
                    c[IDX(i, j, *ldc)] = *beta == 0 ? (T)0 : *beta * c[IDX(i, j, *ldc)];         \
            }                                                                                    \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        gb_require_device();                                                                     \
        void *s = gpublas_dev_stream();                                                          \
        gb_operand ops[2] = {{0}};                                                               \
        if (!scale_only)                                                                         \
            gb_acquire_matrix(&ops[0], a, *lda, nrowa, notr ? *k : *n, sizeof(T), GB_IN, s);     \
        /* only one triangle is written: the other must round-trip unchanged */                  \
        gb_acquire_matrix(&ops[1], c, *ldc, *n, *n, sizeof(T), GB_INOUT, s);                     \
        gb_check(DEVFN(s, *uplo, *trans, *n, *k, *alpha, (const T *)ops[0].dev, *lda, *beta,     \
                       (T *)ops[1].dev, *ldc),                                                   \
                 #PRE "syrk launch");                                                            \
        gb_finish(s, ops, 2);                                                                    \
        GB_COUNT(device_calls, 1);                                                               \
        gb_leave();                                                                              \
    }

DEFINE_SYRK(d, double, "DSYRK ", dsyrk_dev, GB_RK_DSYRK)
DEFINE_SYRK(s, float, "SSYRK ", gpublas_dev_ssyrk, GB_RK_SSYRK)

/* ------------------------------------------------------------------ */
/* SYR2K (netlib DSYR2K; the reference's syr2k.cc:78-146 check, info   */
/* 1 uplo, 2 trans, 3 n, 4 k, 7 lda, 9 ldb, 12 ldc)                    */
/* ------------------------------------------------------------------ */
#define DEFINE_SYR2K(PRE, T, NAME6, DEVFN, RK)                                                   \
    void PRE##syr2k_(const char *uplo, const char *trans, const int *n, const int *k,            \
                     const T *alpha, const T *a, const int *lda, const T *b, const int *ldb,     \
                     const T *beta, T *c, const int *ldc)                                        \
    {                                                                                            \
        gb_enter();                                                                              \
        GB_COUNT(calls, 1);                                                                      \
        const int notr = gb_lsame(*trans, 'N');                                                  \
        const int nrowa = notr ? *n : *k;                                                        \
        const int upper = gb_lsame(*uplo, 'U');                                                  \
        int info = 0;                                                                            \
        if (!upper && !gb_lsame(*uplo, 'L')) info = 1;                                           \
        else if (!notr && !gb_lsame(*trans, 'T') && !gb_lsame(*trans, 'C')) info = 2;            \
        else if (*n < 0) info = 3;                                                               \
        else if (*k < 0) info = 4;                                                               \
        else if (*lda < MAXI(1, nrowa)) info = 7;                                                \
        else if (*ldb < MAXI(1, nrowa)) info = 9;                                                \
        else if (*ldc < MAXI(1, *n)) info = 12;                                                  \
        if (info) {                                                                              \
            gb_xerbla(NAME6, info);                                                              \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        if (*n == 0 || ((*alpha == 0 || *k == 0) && *beta == 1)) {                               \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        {                                                                                        \
            const double es_ = sizeof(T), nn_ = *n, kk_ = *k;                                    \
            if (gb_host_route(RK, 0, 2 * nn_ * nn_ * kk_,                                        \
                              es_ * (2 * nn_ * kk_ + 2 * nn_ * nn_), a, b, c)) {                 \
                typedef void (*fn_t)(const char *, const char *, const int *, const int *,       \
                                     const T *, const T *, const int *, const T *, const int *,  \
                                     const T *, T *, const int *);                               \
                GB_COUNT(fallthrough, 1);                                                        \
                ((fn_t)gb_host_fn(#PRE "syr2k_"))(uplo, trans, n, k, alpha, a, lda, b, ldb,      \
                                                  beta, c, ldc);                                 \
                gb_leave();                                                                      \
                return;                                                                          \
            }                                                                                    \
        }                                                                                        \
        const int scale_only = (*alpha == 0 || *k == 0);                                         \
        if (scale_only && !gb_ptr_tracked(c)) {                                                  \
            for (int j = 0; j < *n; ++j) {                                                       \
                const int i0 = upper ? 0 : j, i1 = upper ? j + 1 : *n;                           \
                for (int i = i0; i < i1; ++i)                                                    \
                    c[IDX(i, j, *ldc)] = *beta == 0 ? (T)0 : *beta * c[IDX(i, j, *ldc)];         \
            }                                                                                    \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        gb_require_device();                                                                     \
        void *s = gpublas_dev_stream();                                                          \
        gb_operand ops[3] = {{0}};                                                               \
        if (!scale_only) {                                                                       \
            gb_acquire_matrix(&ops[0], a, *lda, nrowa, notr ? *k : *n, sizeof(T), GB_IN, s);     \
            gb_acquire_matrix(&ops[1], b, *ldb, nrowa, notr ? *k : *n, sizeof(T), GB_IN, s);     \
        }                                                                                        \
        gb_acquire_matrix(&ops[2], c, *ldc, *n, *n, sizeof(T), GB_INOUT, s);                     \
        gb_check(DEVFN(s, *uplo, *trans, *n, *k, *alpha, (const T *)ops[0].dev, *lda,            \
                       (const T *)ops[1].dev, *ldb, *beta, (T *)ops[2].dev, *ldc),               \
                 #PRE "syr2k launch");                                                           \
        gb_finish(s, ops, 3);                                                                    \
        GB_COUNT(device_calls, 1);                                                               \
        gb_leave();                                                                              \
    }

DEFINE_SYR2K(d, double, "DSYR2K", gpublas_dev_dsyr2k, GB_RK_DSYR2K)
DEFINE_SYR2K(s, float, "SSYR2K", gpublas_dev_ssyr2k, GB_RK_SSYR2K)

/* ------------------------------------------------------------------ */
/* TRMM (netlib DTRMM: info 1 side, 2 uplo, 3 transa, 4 diag, 5 m,     */
/* 6 n, 9 lda, 11 ldb; trmm.cc is not built by the reference)          */
/* ------------------------------------------------------------------ */
#define DEFINE_TRMM(PRE, T, NAME6, DEVFN, RK, HOSTNAME)                                          \
    void PRE##trmm_(const char *side, const char *uplo, const char *transa, const char *diag,    \
                    const int *m, const int *n, const T *alpha, const T *a, const int *lda,      \
                    T *b, const int *ldb)                                                        \
    {                                                                                            \
        gb_enter();                                                                              \
        GB_COUNT(calls, 1);                                                                      \
        const int lside = gb_lsame(*side, 'L');                                                  \
        const int nrowa = lside ? *m : *n;                                                       \
        const int upper = gb_lsame(*uplo, 'U');                                                  \
        int info = 0;                                                                            \
        if (!lside && !gb_lsame(*side, 'R')) info = 1;                                           \
        else if (!upper && !gb_lsame(*uplo, 'L')) info = 2;                                      \
        else if (!gb_lsame(*transa, 'N') && !gb_lsame(*transa, 'T') && !gb_lsame(*transa, 'C'))  \
            info = 3;                                                                            \
        else if (!gb_lsame(*diag, 'U') && !gb_lsame(*diag, 'N')) info = 4;                       \
        else if (*m < 0) info = 5;                                                               \
        else if (*n < 0) info = 6;                                                               \
        else if (*lda < MAXI(1, nrowa)) info = 9;                                                \
        else if (*ldb < MAXI(1, *m)) info = 11;                                                  \
        if (info) {                                                                              \
            gb_xerbla(NAME6, info);                                                              \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        if (*m == 0 || *n == 0) {                                                                \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        {                                                                                        \
            const double es_ = sizeof(T), mm_ = *m, nn_ = *n, na_ = nrowa;                      \
            if (gb_host_route(RK, 0, mm_ * nn_ * na_, es_ * (na_ * na_ + 2 * mm_ * nn_), a, b,   \
                              NULL)) {                                                           \
                typedef void (*fn_t)(const char *, const char *, const char *, const char *,     \
                                     const int *, const int *, const T *, const T *,             \
                                     const int *, T *, const int *);                             \
                GB_COUNT(fallthrough, 1);                                                        \
                ((fn_t)gb_host_fn(HOSTNAME))(side, uplo, transa, diag, m, n, alpha, a, lda, b,   \
                                             ldb);                                               \
                gb_leave();                                                                      \
                return;                                                                          \
            }                                                                                    \
        }                                                                                        \
        if (*alpha == 0 && !gb_ptr_tracked(b)) {                                                 \
            for (int j = 0; j < *n; ++j)                                                         \
                for (int i = 0; i < *m; ++i) b[IDX(i, j, *ldb)] = 0;                             \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        gb_require_device();                                                                     \
        void *s = gpublas_dev_stream();                                                          \
        gb_operand ops[2] = {{0}};                                                               \
        if (*alpha != 0)                                                                         \
            gb_acquire_matrix(&ops[0], a, *lda, nrowa, nrowa, sizeof(T), GB_IN, s);              \
        gb_acquire_matrix(&ops[1], b, *ldb, *m, *n, sizeof(T), *alpha == 0 ? GB_OUT : GB_INOUT,  \
                          s);                                                                    \
        gb_check(DEVFN(s, *side, *uplo, *transa, *diag, *m, *n, *alpha, (const T *)ops[0].dev,   \
                       *lda, (T *)ops[1].dev, *ldb),                                             \
                 #PRE "trmm launch");                                                            \
        gb_finish(s, ops, 2);                                                                    \
        GB_COUNT(device_calls, 1);                                                               \
        gb_leave();                                                                              \
    }

DEFINE_TRMM(d, double, "DTRMM ", gpublas_dev_dtrmm, GB_RK_DTRMM, "dtrmm_")
DEFINE_TRMM(s, float, "STRMM ", gpublas_dev_strmm, GB_RK_STRMM, "strmm_")

/* ------------------------------------------------------------------ */
/* TRSM                                                                */
/* ------------------------------------------------------------------ */
#define DEFINE_TRSM(PRE, T, NAME6, DEVFN, SERVEFN, RK, HOSTNAME)                                      \
    void PRE##trsm_(const char *side, const char *uplo, const char *transa, const char *diag,    \
                    const int *m, const int *n, const T *alpha, const T *a, const int *lda,      \
                    T *b, const int *ldb)                                                        \
    {                                                                                            \
        gb_enter();                                                                              \
        GB_COUNT(calls, 1);                                                                      \
        const int lside = gb_lsame(*side, 'L');                                                  \
        const int nrowa = lside ? *m : *n;                                                       \
        const int upper = gb_lsame(*uplo, 'U');                                                  \
        int info = 0;                                                                            \
        if (!lside && !gb_lsame(*side, 'R')) info = 1;                                           \
        else if (!upper && !gb_lsame(*uplo, 'L')) info = 2;                                      \
        else if (!gb_lsame(*transa, 'N') && !gb_lsame(*transa, 'T') && !gb_lsame(*transa, 'C'))  \
            info = 3;                                                                            \
        else if (!gb_lsame(*diag, 'U') && !gb_lsame(*diag, 'N')) info = 4;                       \
        else if (*m < 0) info = 5;                                                               \
        else if (*n < 0) info = 6;                                                               \
        else if (*lda < MAXI(1, nrowa)) info = 9;                                                \
        else if (*ldb < MAXI(1, *m)) info = 11;                                                  \
        if (info) {                                                                              \
            gb_xerbla(NAME6, info);                                                              \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        if (*m == 0 || *n == 0) {                                                                \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        {                                                                                        \
            const double es_ = sizeof(T), mm_ = *m, nn_ = *n, na_ = nrowa;                      \
            if (gb_host_route(RK, 0, mm_ * nn_ * na_, es_ * (na_ * na_ + 2 * mm_ * nn_), a, b,   \
                              NULL)) {                                                           \
                typedef void (*fn_t)(const char *, const char *, const char *, const char *,     \
                                     const int *, const int *, const T *, const T *,             \
                                     const int *, T *, const int *);                             \
                GB_COUNT(fallthrough, 1);                                                        \
                ((fn_t)gb_host_fn(HOSTNAME))(side, uplo, transa, diag, m, n, alpha, a, lda, b,   \
                                             ldb);                                               \
                gb_leave();                                                                      \
                return;                                                                          \
            }                                                                                    \
        }                                                                                        \
        if (*alpha == 0 && !gb_ptr_tracked(b)) {                                                 \
            for (int j = 0; j < *n; ++j)                                                         \
                for (int i = 0; i < *m; ++i) b[IDX(i, j, *ldb)] = 0;                             \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        if (*alpha != 0 && gb_opt.serve &&                                                       \
            SERVEFN(*side, *uplo, *transa, *diag, *m, *n, *alpha, a, *lda, b, *ldb)) {            \
            GB_COUNT(device_calls, 1);                                                           \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        gb_require_device();                                                                     \
        void *s = gpublas_dev_stream();                                                          \
        gb_operand ops[2] = {{0}};                                                               \
        if (*alpha != 0)                                                                         \
            gb_acquire_matrix(&ops[0], a, *lda, nrowa, nrowa, sizeof(T), GB_IN, s);              \
        gb_acquire_matrix(&ops[1], b, *ldb, *m, *n, sizeof(T), *alpha == 0 ? GB_OUT : GB_INOUT,  \
                          s);                                                                    \
        gb_check(DEVFN(s, *side, *uplo, *transa, *diag, *m, *n, *alpha, (const T *)ops[0].dev,   \
                       *lda, (T *)ops[1].dev, *ldb),                                             \
                 #PRE "trsm launch");                                                            \
        gb_finish(s, ops, 2);                                                                    \
        GB_COUNT(device_calls, 1);                                                               \
        gb_leave();                                                                              \
    }

/* single-precision TRSM has no batch kernel: always the per-call path */
static int no_serve_strsm(char side, char uplo, char transa, char diag, int m, int n, float alpha,
                          const float *a, int lda, float *b, int ldb)
{
    return 0;
}

DEFINE_TRSM(d, double, "DTRSM ", gpublas_dev_dtrsm, gb_serve_dtrsm, GB_RK_DTRSM, "dtrsm_")
DEFINE_TRSM(s, float, "STRSM ", gpublas_dev_strsm, no_serve_strsm, GB_RK_STRSM, "strsm_")

/* ------------------------------------------------------------------ */
/* SYMM                                                                */
/* ------------------------------------------------------------------ */
#define DEFINE_SYMM(PRE, T, NAME6, DEVFN, RK)                                                    \
    void PRE##symm_(const char *side, const char *uplo, const int *m, const int *n,              \
                    const T *alpha, const T *a, const int *lda, const T *b, const int *ldb,      \
                    const T *beta, T *c, const int *ldc)                                         \
    {                                                                                            \
        gb_enter();                                                                              \
        GB_COUNT(calls, 1);                                                                      \
        const int lside = gb_lsame(*side, 'L');                                                  \
        const int nrowa = lside ? *m : *n;                                                       \
        const int upper = gb_lsame(*uplo, 'U');                                                  \
        int info = 0;                                                                            \
        if (!lside && !gb_lsame(*side, 'R')) info = 1;                                           \
        else if (!upper && !gb_lsame(*uplo, 'L')) info = 2;                                      \
        else if (*m < 0) info = 3;                                                               \
        else if (*n < 0) info = 4;                                                               \
        else if (*lda < MAXI(1, nrowa)) info = 7;                                                \
        else if (*ldb < MAXI(1, *m)) info = 9;                                                   \
        else if (*ldc < MAXI(1, *m)) info = 12;                                                  \
        if (info) {                                                                              \
            gb_xerbla(NAME6, info);                                                              \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        if (*m == 0 || *n == 0 || (*alpha == 0 && *beta == 1)) {                                 \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        {                                                                                        \
            const double es_ = sizeof(T), mm_ = *m, nn_ = *n, na_ = nrowa;                      \
            if (gb_host_route(RK, 0, 2 * mm_ * nn_ * na_, es_ * (na_ * na_ + 3 * mm_ * nn_), a,  \
                              b, c)) {                                                           \
                typedef void (*fn_t)(const char *, const char *, const int *, const int *,       \
                                     const T *, const T *, const int *, const T *, const int *,  \
                                     const T *, T *, const int *);                               \
                GB_COUNT(fallthrough, 1);                                                        \
                ((fn_t)gb_host_fn(#PRE "symm_"))(side, uplo, m, n, alpha, a, lda, b, ldb, beta,  \
                                                 c, ldc);                                        \
                gb_leave();                                                                      \
                return;                                                                          \
            }                                                                                    \
        }                                                                                        \
        if (*alpha == 0 && !gb_ptr_tracked(c)) {                                                 \
            for (int j = 0; j < *n; ++j)                                                         \
                for (int i = 0; i < *m; ++i)                                                     \
                    c[IDX(i, j, *ldc)] = *beta == 0 ? (T)0 : *beta * c[IDX(i, j, *ldc)];         \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        gb_require_device();                                                                     \
        void *s = gpublas_dev_stream();                                                          \
        gb_operand ops[3] = {{0}};                                                               \
        if (*alpha != 0) {                                                                       \
            gb_acquire_matrix(&ops[0], a, *lda, nrowa, nrowa, sizeof(T), GB_IN, s);              \
            gb_acquire_matrix(&ops[1], b, *ldb, *m, *n, sizeof(T), GB_IN, s);                    \
        }                                                                                        \
        gb_acquire_matrix(&ops[2], c, *ldc, *m, *n, sizeof(T), *beta == 0 ? GB_OUT : GB_INOUT,  \
                          s);                                                                    \
        gb_check(DEVFN(s, *side, *uplo, *m, *n, *alpha, (const T *)ops[0].dev, *lda,             \
                       (const T *)ops[1].dev, *ldb, *beta, (T *)ops[2].dev, *ldc),               \
                 #PRE "symm launch");                                                            \
        gb_finish(s, ops, 3);                                                                    \
        GB_COUNT(device_calls, 1);                                                               \
        gb_leave();                                                                              \
    }

DEFINE_SYMM(d, double, "DSYMM ", gpublas_dev_dsymm, GB_RK_DSYMM)
DEFINE_SYMM(s, float, "SSYMM ", gpublas_dev_ssymm, GB_RK_SSYMM)

/* ------------------------------------------------------------------ */
/* complex SYRK / HERK / SYR2K / HER2K / SYMM / HEMM / TRMM / TRSM       */
/* (syrk.cc, herk.cc, syr2k.cc, her2k.cc, symm.cc, hemm.cc, trsm.cc of   */
/* the reference; netlib semantics; trmm.cc is not built there).         */
/* Scalars are (re, im) pairs, except the real alpha / beta of HERK and  */
/* the real beta of HER2K.  A call that only scales C (alpha = 0 or k =  */
/* 0) runs the scaling kernel on the device.                             */
/* ------------------------------------------------------------------ */
#define CZERO(p) ((p)[0] == 0 && (p)[1] == 0)
#define CONE(p) ((p)[0] == 1 && (p)[1] == 0)

#define GB_CPLX_HOST(PRE, NAME, ARGS_T, ARGS)                                                    \
    do {                                                                                         \
        typedef void(*fn_t) ARGS_T;                                                              \
        fn_t f = (fn_t)gb_host_fn(#PRE #NAME "_");                                               \
        GB_COUNT(fallthrough, 1);                                                                \
        f ARGS;                                                                                  \
        gb_leave();                                                                              \
        return;                                                                                  \
    } while (0)

/* RK: SYRK-like (one operand) or SYR2K-like (two); HERM: real alpha/beta (HERK) or real beta
 * (HER2K) */
#define DEFINE_CRK(PRE, NAME, R, NAME6, HERM, TWO, RK, AL_T, BE_T, DEVCALL)                       \
    void PRE##NAME##_(const char *uplo, const char *trans, const int *n, const int *k,           \
                      const AL_T *alpha, const R *a, const int *lda TWO##_PARAM(R),              \
                      const BE_T *beta, R *c, const int *ldc)                                    \
    {                                                                                            \
        gb_enter();                                                                              \
        GB_COUNT(calls, 1);                                                                      \
        const int upper = gb_lsame(*uplo, 'U'), nt = gb_lsame(*trans, 'N');                     \
        const int nrowa = nt ? *n : *k;                                                          \
        int info = 0;                                                                            \
        if (!upper && !gb_lsame(*uplo, 'L')) info = 1;                                           \
        else if (!nt && !gb_lsame(*trans, HERM ? 'C' : 'T')) info = 2;                           \
        else if (*n < 0) info = 3;                                                               \
        else if (*k < 0) info = 4;                                                               \
        else if (*lda < MAXI(1, nrowa)) info = 7;                                                \
        TWO##_LDB_CHECK                                                                          \
        else if (*ldc < MAXI(1, *n)) info = TWO##_LDC_INFO;                                      \
        if (info) {                                                                              \
            gb_xerbla(NAME6, info);                                                              \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        const int a0 = AL_T##_ZERO(alpha), b1 = BE_T##_ONE(beta);                                \
        if (*n == 0 || ((a0 || *k == 0) && b1)) {                                                \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        const int cut = gb_opt.cutoff;                                                           \
        const double es_ = 2 * sizeof(R), nn_ = *n, kk_ = *k;                                    \
        const int scale_only = a0 || *k == 0;                                                    \
        if (gb_host_route(RK, cut > 0 && (*lda < cut || *ldc < cut), 4 * (TWO##_F) * nn_ * nn_ * kk_, \
                          es_ * ((TWO##_F) * nn_ * kk_ + 2 * nn_ * nn_), a, TWO##_PB, c))        \
            GB_CPLX_HOST(PRE, NAME,                                                              \
                         (const char *, const char *, const int *, const int *, const AL_T *,    \
                          const R *, const int *TWO##_PT(R), const BE_T *, R *, const int *),    \
                         (uplo, trans, n, k, alpha, a, lda TWO##_PA, beta, c, ldc));             \
        gb_require_device();                                                                     \
        void *s = gpublas_dev_stream();                                                          \
        gb_operand ops[3] = {{0}};                                                               \
        if (!scale_only) {                                                                       \
            gb_acquire_matrix(&ops[0], a, *lda, nrowa, nt ? *k : *n, 2 * sizeof(R), GB_IN, s);   \
            TWO##_ACQ(R)                                                                         \
        }                                                                                        \
        gb_acquire_matrix(&ops[2], c, *ldc, *n, *n, 2 * sizeof(R), GB_INOUT, s);                 \
        gb_check(DEVCALL, #PRE #NAME " launch");                                                 \
        gb_finish(s, ops, 3);                                                                    \
        GB_COUNT(device_calls, 1);                                                               \
        gb_leave();                                                                              \
    }
/* the pieces that differ between the one- and two-operand forms */
#define ONE_PARAM(R)
#define TWO_PARAM(R) , const R *b, const int *ldb
#define ONE_PT(R)
#define TWO_PT(R) , const R *, const int *
#define ONE_PA
#define TWO_PA , b, ldb
#define ONE_PB NULL
#define TWO_PB b
#define ONE_LDB_CHECK
#define TWO_LDB_CHECK else if (*ldb < MAXI(1, nrowa)) info = 9;
#define ONE_LDC_INFO 10
#define TWO_LDC_INFO 12
#define ONE_F 1
#define TWO_F 2
#define ONE_ACQ(R)
#define TWO_ACQ(R) gb_acquire_matrix(&ops[1], b, *ldb, nrowa, nt ? *k : *n, 2 * sizeof(R), GB_IN, s);
#define double_ZERO(p) (*(p) == 0)
#define double_ONE(p) (*(p) == 1)
#define float_ZERO(p) (*(p) == 0)
#define float_ONE(p) (*(p) == 1)
/* complex scalars are spelled CD / CF here: (re, im) pairs */
typedef double CD;
typedef float CF;
#define CD_ZERO(p) CZERO(p)
#define CD_ONE(p) CONE(p)
#define CF_ZERO(p) CZERO(p)
#define CF_ONE(p) CONE(p)

#define DEV_A ((const R_ *)ops[0].dev)
#define DEV_B ((const R_ *)ops[1].dev)
#define DEV_C ((R_ *)ops[2].dev)
#define R_ double
DEFINE_CRK(z, syrk, double, "ZSYRK ", 0, ONE, GB_RK_DSYRK, CD, CD,
           gpublas_dev_zsyrk(s, *uplo, *trans, *n, *k, alpha[0], alpha[1], DEV_A, *lda, beta[0],
                             beta[1], DEV_C, *ldc))
DEFINE_CRK(z, herk, double, "ZHERK ", 1, ONE, GB_RK_DSYRK, double, double,
           gpublas_dev_zherk(s, *uplo, *trans, *n, *k, *alpha, DEV_A, *lda, *beta, DEV_C, *ldc))
DEFINE_CRK(z, syr2k, double, "ZSYR2K", 0, TWO, GB_RK_DSYR2K, CD, CD,
           gpublas_dev_zsyr2k(s, *uplo, *trans, *n, *k, alpha[0], alpha[1], DEV_A, *lda, DEV_B,
                              *ldb, beta[0], beta[1], DEV_C, *ldc))
DEFINE_CRK(z, her2k, double, "ZHER2K", 1, TWO, GB_RK_DSYR2K, CD, double,
           gpublas_dev_zher2k(s, *uplo, *trans, *n, *k, alpha[0], alpha[1], DEV_A, *lda, DEV_B,
                              *ldb, *beta, DEV_C, *ldc))
#undef R_
#define R_ float
DEFINE_CRK(c, syrk, float, "CSYRK ", 0, ONE, GB_RK_SSYRK, CF, CF,
           gpublas_dev_csyrk(s, *uplo, *trans, *n, *k, alpha[0], alpha[1], DEV_A, *lda, beta[0],
                             beta[1], DEV_C, *ldc))
DEFINE_CRK(c, herk, float, "CHERK ", 1, ONE, GB_RK_SSYRK, float, float,
           gpublas_dev_cherk(s, *uplo, *trans, *n, *k, *alpha, DEV_A, *lda, *beta, DEV_C, *ldc))
DEFINE_CRK(c, syr2k, float, "CSYR2K", 0, TWO, GB_RK_SSYR2K, CF, CF,
           gpublas_dev_csyr2k(s, *uplo, *trans, *n, *k, alpha[0], alpha[1], DEV_A, *lda, DEV_B,
                              *ldb, beta[0], beta[1], DEV_C, *ldc))
DEFINE_CRK(c, her2k, float, "CHER2K", 1, TWO, GB_RK_SSYR2K, CF, float,
           gpublas_dev_cher2k(s, *uplo, *trans, *n, *k, alpha[0], alpha[1], DEV_A, *lda, DEV_B,
                              *ldb, *beta, DEV_C, *ldc))
#undef R_

#define DEFINE_CSYMM(PRE, NAME, R, NAME6, HERM, RK, DEVFN)                                        \
    void PRE##NAME##_(const char *side, const char *uplo, const int *m, const int *n,            \
                      const R *alpha, const R *a, const int *lda, const R *b, const int *ldb,    \
                      const R *beta, R *c, const int *ldc)                                       \
    {                                                                                            \
        gb_enter();                                                                              \
        GB_COUNT(calls, 1);                                                                      \
        const int left = gb_lsame(*side, 'L'), upper = gb_lsame(*uplo, 'U');                     \
        const int na = left ? *m : *n;                                                           \
        int info = 0;                                                                            \
        if (!left && !gb_lsame(*side, 'R')) info = 1;                                            \
        else if (!upper && !gb_lsame(*uplo, 'L')) info = 2;                                      \
        else if (*m < 0) info = 3;                                                               \
        else if (*n < 0) info = 4;                                                               \
        else if (*lda < MAXI(1, na)) info = 7;                                                   \
        else if (*ldb < MAXI(1, *m)) info = 9;                                                   \
        else if (*ldc < MAXI(1, *m)) info = 12;                                                  \
        if (info) {                                                                              \
            gb_xerbla(NAME6, info);                                                              \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        if (*m == 0 || *n == 0 || (CZERO(alpha) && CONE(beta))) {                                \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        const int cut = gb_opt.cutoff;                                                           \
        const double es_ = 2 * sizeof(R), mm_ = *m, nn_ = *n, aa_ = na;                          \
        if (gb_host_route(RK, cut > 0 && (*lda < cut || *ldb < cut || *ldc < cut),             \
                          8 * mm_ * nn_ * aa_, es_ * (aa_ * aa_ + 3 * mm_ * nn_), a, b, c))      \
            GB_CPLX_HOST(PRE, NAME,                                                              \
                         (const char *, const char *, const int *, const int *, const R *,       \
                          const R *, const int *, const R *, const int *, const R *, R *,        \
                          const int *),                                                          \
                         (side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc));               \
        gb_require_device();                                                                     \
        void *s = gpublas_dev_stream();                                                          \
        gb_operand ops[3] = {{0}};                                                               \
        if (!CZERO(alpha)) {                                                                     \
            gb_acquire_matrix(&ops[0], a, *lda, na, na, 2 * sizeof(R), GB_IN, s);                \
            gb_acquire_matrix(&ops[1], b, *ldb, *m, *n, 2 * sizeof(R), GB_IN, s);                \
        }                                                                                        \
        gb_acquire_matrix(&ops[2], c, *ldc, *m, *n, 2 * sizeof(R),                               \
                          CZERO(beta) ? GB_OUT : GB_INOUT, s);                                   \
        gb_check(DEVFN(s, HERM, *side, *uplo, *m, *n, alpha[0], alpha[1], (const R *)ops[0].dev, \
                       *lda, (const R *)ops[1].dev, *ldb, beta[0], beta[1], (R *)ops[2].dev,     \
                       *ldc),                                                                    \
                 #PRE #NAME " launch");                                                          \
        gb_finish(s, ops, 3);                                                                    \
        GB_COUNT(device_calls, 1);                                                               \
        gb_leave();                                                                              \
    }
DEFINE_CSYMM(z, symm, double, "ZSYMM ", 0, GB_RK_DSYMM, gpublas_dev_zsymm)
DEFINE_CSYMM(z, hemm, double, "ZHEMM ", 1, GB_RK_DSYMM, gpublas_dev_zsymm)
DEFINE_CSYMM(c, symm, float, "CSYMM ", 0, GB_RK_SSYMM, gpublas_dev_csymm)
DEFINE_CSYMM(c, hemm, float, "CHEMM ", 1, GB_RK_SSYMM, gpublas_dev_csymm)

#define DEFINE_CTRI(PRE, NAME, R, NAME6, RK, DEVFN)                                               \
    void PRE##NAME##_(const char *side, const char *uplo, const char *transa, const char *diag,  \
                      const int *m, const int *n, const R *alpha, const R *a, const int *lda,    \
                      R *b, const int *ldb)                                                      \
    {                                                                                            \
        gb_enter();                                                                              \
        GB_COUNT(calls, 1);                                                                      \
        const int left = gb_lsame(*side, 'L'), upper = gb_lsame(*uplo, 'U');                     \
        const int na = left ? *m : *n;                                                           \
        int info = 0;                                                                            \
        if (!left && !gb_lsame(*side, 'R')) info = 1;                                            \
        else if (!upper && !gb_lsame(*uplo, 'L')) info = 2;                                      \
        else if (!gb_lsame(*transa, 'N') && !gb_lsame(*transa, 'T') && !gb_lsame(*transa, 'C')) \
            info = 3;                                                                            \
        else if (!gb_lsame(*diag, 'U') && !gb_lsame(*diag, 'N')) info = 4;                       \
        else if (*m < 0) info = 5;                                                               \
        else if (*n < 0) info = 6;                                                               \
        else if (*lda < MAXI(1, na)) info = 9;                                                   \
        else if (*ldb < MAXI(1, *m)) info = 11;                                                  \
        if (info) {                                                                              \
            gb_xerbla(NAME6, info);                                                              \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        if (*m == 0 || *n == 0) {                                                                \
            gb_leave();                                                                          \
            return;                                                                              \
        }                                                                                        \
        const int cut = gb_opt.cutoff;                                                           \
        const double es_ = 2 * sizeof(R), mm_ = *m, nn_ = *n, aa_ = na;                          \
        if (gb_host_route(RK, cut > 0 && (*lda < cut || *ldb < cut), 4 * mm_ * nn_ * aa_,       \
                          es_ * (aa_ * aa_ / 2 + 2 * mm_ * nn_), a, b, NULL))                    \
            GB_CPLX_HOST(PRE, NAME,                                                              \
                         (const char *, const char *, const char *, const char *, const int *,   \
                          const int *, const R *, const R *, const int *, R *, const int *),     \
                         (side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb));               \
        gb_require_device();                                                                     \
        void *s = gpublas_dev_stream();                                                          \
        gb_operand ops[2] = {{0}};                                                               \
        if (!CZERO(alpha)) gb_acquire_matrix(&ops[0], a, *lda, na, na, 2 * sizeof(R), GB_IN, s); \
        gb_acquire_matrix(&ops[1], b, *ldb, *m, *n, 2 * sizeof(R),                               \
                          CZERO(alpha) ? GB_OUT : GB_INOUT, s);                                  \
        gb_check(DEVFN(s, *side, *uplo, *transa, *diag, *m, *n, alpha[0], alpha[1],              \
                       (const R *)ops[0].dev, *lda, (R *)ops[1].dev, *ldb),                      \
                 #PRE #NAME " launch");                                                          \
        gb_finish(s, ops, 2);                                                                    \
        GB_COUNT(device_calls, 1);                                                               \
        gb_leave();                                                                              \
    }
DEFINE_CTRI(z, trsm, double, "ZTRSM ", GB_RK_DTRSM, gpublas_dev_ztrsm)
DEFINE_CTRI(c, trsm, float, "CTRSM ", GB_RK_STRSM, gpublas_dev_ctrsm)
DEFINE_CTRI(z, trmm, double, "ZTRMM ", GB_RK_DTRMM, gpublas_dev_ztrmm)
DEFINE_CTRI(c, trmm, float, "CTRMM ", GB_RK_STRMM, gpublas_dev_ctrmm)

/* Library-internal names of the entry points above, bound inside the library
 * (the crossover measurement calls them; a process may hold another BLAS
 * whose exported dgemm_ would win an ordinary symbol lookup). */
#define GB_SELF(f) extern __typeof__(f) gb_self_##f __attribute__((alias(#f), visibility("hidden")));
GB_SELF(dgemm_)
GB_SELF(sgemm_)
GB_SELF(dsyrk_)
GB_SELF(ssyrk_)
GB_SELF(dsyr2k_)
GB_SELF(ssyr2k_)
GB_SELF(dsymm_)
GB_SELF(ssymm_)
GB_SELF(dtrsm_)
GB_SELF(strsm_)
GB_SELF(dtrmm_)
GB_SELF(strmm_)
