/*
 * cblas.c -- CBLAS twins of the hot-path routines (cblas.h:57-822 of the
 * reference; enum values cblas.h:21-25).
 *
 * Column-major calls forward to the Fortran entries; row-major calls are the
 * same column-major problem on the transposed operands (swap the operand
 * roles and flip trans / uplo / side), which is exact -- no data moves.
 * Interposing cblas_* matters because a host CBLAS does not route its cblas_
 * functions through the Fortran symbols.
 *
 * Errors are reported through xerbla with the CBLAS parameter position
 * (layout = 1), the convention of the netlib CBLAS and OpenBLAS front ends.
 */
#include "internal.h"

static void cbl_err(const char *name6, int pos) { gb_xerbla(name6, pos); }

static char tr(CBLAS_TRANSPOSE t) { return t == CblasNoTrans ? 'N' : t == CblasTrans ? 'T' : 'C'; }
static char trflip(CBLAS_TRANSPOSE t) { return t == CblasNoTrans ? 'T' : 'N'; }
static int tr_ok(CBLAS_TRANSPOSE t)
{
    return t == CblasNoTrans || t == CblasTrans || t == CblasConjTrans;
}
static int up_ok(CBLAS_UPLO u) { return u == CblasUpper || u == CblasLower; }
static char up(CBLAS_UPLO u) { return u == CblasUpper ? 'U' : 'L'; }
static char upflip(CBLAS_UPLO u) { return u == CblasUpper ? 'L' : 'U'; }
static int side_ok(CBLAS_SIDE s) { return s == CblasLeft || s == CblasRight; }
static char sd(CBLAS_SIDE s) { return s == CblasLeft ? 'L' : 'R'; }
static char sdflip(CBLAS_SIDE s) { return s == CblasLeft ? 'R' : 'L'; }
static int diag_ok(CBLAS_DIAG d) { return d == CblasUnit || d == CblasNonUnit; }
static char dg(CBLAS_DIAG d) { return d == CblasUnit ? 'U' : 'N'; }
static int lay_ok(CBLAS_LAYOUT l) { return l == CblasRowMajor || l == CblasColMajor; }

/* Map the Fortran INFO of the forwarded call to the CBLAS position.
 * fmap[i] = CBLAS position for Fortran INFO i (0 = none). */
static GB_TLS const int *t_map = 0;
static GB_TLS int t_map_n = 0;
extern GB_TLS int gb_cblas_active;
GB_TLS int gb_cblas_active = 0;

int gb_cblas_translate(int info)
{
    if (!gb_cblas_active || !t_map || info <= 0 || info >= t_map_n) return info;
    return t_map[info] ? t_map[info] : info;
}

#define WITH_MAP(map, call)                                                                      \
    do {                                                                                         \
        t_map = map;                                                                             \
        t_map_n = (int)(sizeof(map) / sizeof(map[0]));                                           \
        gb_cblas_active = 1;                                                                     \
        call;                                                                                    \
        gb_cblas_active = 0;                                                                     \
        t_map = 0;                                                                               \
    } while (0)

/* ---- GEMM: cblas positions 1 layout 2 ta 3 tb 4 m 5 n 6 k 7 alpha 8 A 9 lda
 *      10 B 11 ldb 12 beta 13 C 14 ldc ---- */
static const int gemm_col[14] = {0, 2, 3, 4, 5, 6, 0, 0, 9, 0, 11, 0, 0, 14};
static const int gemm_row[14] = {0, 3, 2, 5, 4, 6, 0, 0, 11, 0, 9, 0, 0, 14};

#define DEFINE_CBLAS_GEMM(PRE, T, NAME6)                                                         \
    void cblas_##PRE##gemm(CBLAS_LAYOUT layout, CBLAS_TRANSPOSE ta, CBLAS_TRANSPOSE tb, int m,    \
                           int n, int k, T alpha, const T *a, int lda, const T *b, int ldb,      \
                           T beta, T *c, int ldc)                                                \
    {                                                                                            \
        if (!lay_ok(layout)) { cbl_err(NAME6, 1); return; }                                 \
        if (!tr_ok(ta)) { cbl_err(NAME6, 2); return; }                                                \
        if (!tr_ok(tb)) { cbl_err(NAME6, 3); return; }                                                \
        char cta = tr(ta), ctb = tr(tb);                                                         \
        if (layout == CblasColMajor)                                                             \
            WITH_MAP(gemm_col, PRE##gemm_(&cta, &ctb, &m, &n, &k, &alpha, a, &lda, b, &ldb,      \
                                          &beta, c, &ldc));                                      \
        else                                                                                     \
            WITH_MAP(gemm_row, PRE##gemm_(&ctb, &cta, &n, &m, &k, &alpha, b, &ldb, a, &lda,      \
                                          &beta, c, &ldc));                                      \
    }
DEFINE_CBLAS_GEMM(d, double, "DGEMM ")
DEFINE_CBLAS_GEMM(s, float, "SGEMM ")

/* complex twins: alpha and beta by address, (re, im) pairs */
#define DEFINE_CBLAS_CGEMM(PRE, R, NAME6)                                                        \
    void cblas_##PRE##gemm(CBLAS_LAYOUT layout, CBLAS_TRANSPOSE ta, CBLAS_TRANSPOSE tb, int m,    \
                           int n, int k, const void *alpha, const void *a, int lda,              \
                           const void *b, int ldb, const void *beta, void *c, int ldc)           \
    {                                                                                            \
        if (!lay_ok(layout)) { cbl_err(NAME6, 1); return; }                                      \
        if (!tr_ok(ta)) { cbl_err(NAME6, 2); return; }                                           \
        if (!tr_ok(tb)) { cbl_err(NAME6, 3); return; }                                           \
        char cta = tr(ta), ctb = tr(tb);                                                         \
        const R *al = (const R *)alpha, *be = (const R *)beta;                                   \
        if (layout == CblasColMajor)                                                             \
            WITH_MAP(gemm_col, PRE##gemm_(&cta, &ctb, &m, &n, &k, al, (const R *)a, &lda,        \
                                          (const R *)b, &ldb, be, (R *)c, &ldc));                \
        else                                                                                     \
            WITH_MAP(gemm_row, PRE##gemm_(&ctb, &cta, &n, &m, &k, al, (const R *)b, &ldb,        \
                                          (const R *)a, &lda, be, (R *)c, &ldc));                \
    }
DEFINE_CBLAS_CGEMM(z, double, "ZGEMM ")
DEFINE_CBLAS_CGEMM(c, float, "CGEMM ")

/* ---- SYRK: 1 layout 2 uplo 3 trans 4 n 5 k 6 alpha 7 A 8 lda 9 beta 10 C 11 ldc ---- */
static const int syrk_map[11] = {0, 2, 3, 4, 5, 0, 0, 8, 0, 0, 11};

#define DEFINE_CBLAS_SYRK(PRE, T, NAME6)                                                         \
    void cblas_##PRE##syrk(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, int n,   \
                           int k, T alpha, const T *a, int lda, T beta, T *c, int ldc)           \
    {                                                                                            \
        if (!lay_ok(layout)) { cbl_err(NAME6, 1); return; }                                      \
        if (!up_ok(uplo)) { cbl_err(NAME6, 2); return; }                                         \
        if (!tr_ok(trans)) { cbl_err(NAME6, 3); return; }                                        \
        char u = layout == CblasColMajor ? up(uplo) : upflip(uplo);                              \
        char t = layout == CblasColMajor ? tr(trans) : trflip(trans);                            \
        WITH_MAP(syrk_map, PRE##syrk_(&u, &t, &n, &k, &alpha, a, &lda, &beta, c, &ldc));         \
    }
DEFINE_CBLAS_SYRK(d, double, "DSYRK ")
DEFINE_CBLAS_SYRK(s, float, "SSYRK ")

/* ---- TRSM / TRMM: 1 layout 2 side 3 uplo 4 ta 5 diag 6 m 7 n 8 alpha 9 A 10 lda 11 B
 *      12 ldb (row major: the transposed problem, side and uplo flipped) ---- */
static const int trsm_col[12] = {0, 2, 3, 4, 5, 6, 7, 0, 0, 10, 0, 12};
static const int trsm_row[12] = {0, 2, 3, 4, 5, 7, 6, 0, 0, 10, 0, 12};

#define DEFINE_CBLAS_TRIANGULAR(PRE, OP, T, NAME6)                                               \
    void cblas_##PRE##OP(CBLAS_LAYOUT layout, CBLAS_SIDE side, CBLAS_UPLO uplo,                  \
                         CBLAS_TRANSPOSE ta, CBLAS_DIAG diag, int m, int n, T alpha, const T *a, \
                         int lda, T *b, int ldb)                                                 \
    {                                                                                            \
        if (!lay_ok(layout)) { cbl_err(NAME6, 1); return; }                                      \
        if (!side_ok(side)) { cbl_err(NAME6, 2); return; }                                       \
        if (!up_ok(uplo)) { cbl_err(NAME6, 3); return; }                                         \
        if (!tr_ok(ta)) { cbl_err(NAME6, 4); return; }                                           \
        if (!diag_ok(diag)) { cbl_err(NAME6, 5); return; }                                       \
        char t = tr(ta), d = dg(diag);                                                           \
        if (layout == CblasColMajor) {                                                           \
            char s = sd(side), u = up(uplo);                                                     \
            WITH_MAP(trsm_col, PRE##OP##_(&s, &u, &t, &d, &m, &n, &alpha, a, &lda, b, &ldb));    \
        } else {                                                                                 \
            char s = sdflip(side), u = upflip(uplo);                                             \
            WITH_MAP(trsm_row, PRE##OP##_(&s, &u, &t, &d, &n, &m, &alpha, a, &lda, b, &ldb));    \
        }                                                                                        \
    }
DEFINE_CBLAS_TRIANGULAR(d, trsm, double, "DTRSM ")
DEFINE_CBLAS_TRIANGULAR(s, trsm, float, "STRSM ")
DEFINE_CBLAS_TRIANGULAR(d, trmm, double, "DTRMM ")
DEFINE_CBLAS_TRIANGULAR(s, trmm, float, "STRMM ")

/* ---- SYMM: 1 layout 2 side 3 uplo 4 m 5 n 6 alpha 7 A 8 lda 9 B 10 ldb 11 beta 12 C 13 ldc ---- */
static const int symm_col[13] = {0, 2, 3, 4, 5, 0, 0, 8, 0, 10, 0, 0, 13};
static const int symm_row[13] = {0, 2, 3, 5, 4, 0, 0, 8, 0, 10, 0, 0, 13};

#define DEFINE_CBLAS_SYMM(PRE, T, NAME6)                                                         \
    void cblas_##PRE##symm(CBLAS_LAYOUT layout, CBLAS_SIDE side, CBLAS_UPLO uplo, int m, int n,  \
                           T alpha, const T *a, int lda, const T *b, int ldb, T beta, T *c,      \
                           int ldc)                                                              \
    {                                                                                            \
        if (!lay_ok(layout)) { cbl_err(NAME6, 1); return; }                                      \
        if (!side_ok(side)) { cbl_err(NAME6, 2); return; }                                       \
        if (!up_ok(uplo)) { cbl_err(NAME6, 3); return; }                                         \
        if (layout == CblasColMajor) {                                                           \
            char s = sd(side), u = up(uplo);                                                     \
            WITH_MAP(symm_col,                                                                   \
                     PRE##symm_(&s, &u, &m, &n, &alpha, a, &lda, b, &ldb, &beta, c, &ldc));      \
        } else {                                                                                 \
            char s = sdflip(side), u = upflip(uplo);                                             \
            WITH_MAP(symm_row,                                                                   \
                     PRE##symm_(&s, &u, &n, &m, &alpha, a, &lda, b, &ldb, &beta, c, &ldc));      \
        }                                                                                        \
    }
DEFINE_CBLAS_SYMM(d, double, "DSYMM ")
DEFINE_CBLAS_SYMM(s, float, "SSYMM ")

/* ---- SYR2K: 1 layout 2 uplo 3 trans 4 n 5 k 6 alpha 7 A 8 lda 9 B 10 ldb 11 beta 12 C
 *      13 ldc (row major: the same update of C^T, uplo and trans flipped) ---- */
static const int syr2k_map[13] = {0, 2, 3, 4, 5, 0, 0, 8, 0, 10, 0, 0, 13};

#define DEFINE_CBLAS_SYR2K(PRE, T, NAME6)                                                        \
    void cblas_##PRE##syr2k(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, int n,  \
                            int k, T alpha, const T *a, int lda, const T *b, int ldb, T beta,    \
                            T *c, int ldc)                                                       \
    {                                                                                            \
        if (!lay_ok(layout)) { cbl_err(NAME6, 1); return; }                                      \
        if (!up_ok(uplo)) { cbl_err(NAME6, 2); return; }                                         \
        if (!tr_ok(trans)) { cbl_err(NAME6, 3); return; }                                        \
        char u = layout == CblasColMajor ? up(uplo) : upflip(uplo);                              \
        char t = layout == CblasColMajor ? tr(trans) : trflip(trans);                            \
        WITH_MAP(syr2k_map,                                                                      \
                 PRE##syr2k_(&u, &t, &n, &k, &alpha, a, &lda, b, &ldb, &beta, c, &ldc));         \
    }
DEFINE_CBLAS_SYR2K(d, double, "DSYR2K")
DEFINE_CBLAS_SYR2K(s, float, "SSYR2K")

/* ---- complex twins of SYRK / HERK / SYR2K / HER2K / SYMM / HEMM / TRMM / TRSM: complex
 *      scalars by address; row major is the column-major problem on the transposed operands:
 *      uplo flipped and trans N <-> T (N <-> C for the Hermitian forms; HER2K then takes
 *      conj(alpha), since C^T = conj(alpha) conj(A) B^T + alpha conj(B) A^T), side flipped and
 *      m / n swapped for SYMM / HEMM / TRMM / TRSM.  SYRK / SYR2K accept no ConjTrans,
 *      HERK / HER2K no Trans (netlib CBLAS). ---- */
static char trflipc(CBLAS_TRANSPOSE t) { return t == CblasNoTrans ? 'C' : 'N'; }
static int tr_ok_sym(CBLAS_TRANSPOSE t) { return t == CblasNoTrans || t == CblasTrans; }
static int tr_ok_herm(CBLAS_TRANSPOSE t) { return t == CblasNoTrans || t == CblasConjTrans; }

#define DEFINE_CBLAS_CRK(PRE, NAME, R, NAME6, HERM, AT, BT, AARG, BARG)                          \
    void cblas_##PRE##NAME(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, int n,   \
                           int k, AT alpha, const void *a, int lda, BT beta, void *c, int ldc)   \
    {                                                                                            \
        if (!lay_ok(layout)) { cbl_err(NAME6, 1); return; }                                      \
        if (!up_ok(uplo)) { cbl_err(NAME6, 2); return; }                                         \
        if (HERM ? !tr_ok_herm(trans) : !tr_ok_sym(trans)) { cbl_err(NAME6, 3); return; }         \
        char u = layout == CblasColMajor ? up(uplo) : upflip(uplo);                              \
        char t = layout == CblasColMajor ? tr(trans) : (HERM ? trflipc(trans) : trflip(trans));  \
        WITH_MAP(syrk_map, PRE##NAME##_(&u, &t, &n, &k, AARG, (const R *)a, &lda, BARG, (R *)c,  \
                                        &ldc));                                                  \
    }
DEFINE_CBLAS_CRK(z, syrk, double, "ZSYRK ", 0, const void *, const void *, (const double *)alpha,
                 (const double *)beta)
DEFINE_CBLAS_CRK(c, syrk, float, "CSYRK ", 0, const void *, const void *, (const float *)alpha,
                 (const float *)beta)
DEFINE_CBLAS_CRK(z, herk, double, "ZHERK ", 1, double, double, &alpha, &beta)
DEFINE_CBLAS_CRK(c, herk, float, "CHERK ", 1, float, float, &alpha, &beta)

#define DEFINE_CBLAS_CR2K(PRE, NAME, R, NAME6, HERM, BT, BARG)                                    \
    void cblas_##PRE##NAME(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, int n,   \
                           int k, const void *alpha, const void *a, int lda, const void *b,      \
                           int ldb, BT beta, void *c, int ldc)                                   \
    {                                                                                            \
        if (!lay_ok(layout)) { cbl_err(NAME6, 1); return; }                                      \
        if (!up_ok(uplo)) { cbl_err(NAME6, 2); return; }                                         \
        if (HERM ? !tr_ok_herm(trans) : !tr_ok_sym(trans)) { cbl_err(NAME6, 3); return; }         \
        const int row = layout == CblasRowMajor;                                                 \
        char u = row ? upflip(uplo) : up(uplo);                                                  \
        char t = !row ? tr(trans) : (HERM ? trflipc(trans) : trflip(trans));                     \
        const R *al = (const R *)alpha;                                                          \
        R alc[2] = {al[0], HERM && row ? -al[1] : al[1]};                                        \
        WITH_MAP(syr2k_map, PRE##NAME##_(&u, &t, &n, &k, alc, (const R *)a, &lda, (const R *)b,  \
                                         &ldb, BARG, (R *)c, &ldc));                             \
    }
DEFINE_CBLAS_CR2K(z, syr2k, double, "ZSYR2K", 0, const void *, (const double *)beta)
DEFINE_CBLAS_CR2K(c, syr2k, float, "CSYR2K", 0, const void *, (const float *)beta)
DEFINE_CBLAS_CR2K(z, her2k, double, "ZHER2K", 1, double, &beta)
DEFINE_CBLAS_CR2K(c, her2k, float, "CHER2K", 1, float, &beta)

#define DEFINE_CBLAS_CSYMM(PRE, NAME, R, NAME6)                                                  \
    void cblas_##PRE##NAME(CBLAS_LAYOUT layout, CBLAS_SIDE side, CBLAS_UPLO uplo, int m, int n,  \
                           const void *alpha, const void *a, int lda, const void *b, int ldb,    \
                           const void *beta, void *c, int ldc)                                   \
    {                                                                                            \
        if (!lay_ok(layout)) { cbl_err(NAME6, 1); return; }                                      \
        if (!side_ok(side)) { cbl_err(NAME6, 2); return; }                                       \
        if (!up_ok(uplo)) { cbl_err(NAME6, 3); return; }                                         \
        const R *al = (const R *)alpha, *be = (const R *)beta;                                   \
        if (layout == CblasColMajor) {                                                           \
            char s = sd(side), u = up(uplo);                                                     \
            WITH_MAP(symm_col, PRE##NAME##_(&s, &u, &m, &n, al, (const R *)a, &lda,              \
                                            (const R *)b, &ldb, be, (R *)c, &ldc));              \
        } else {                                                                                 \
            char s = sdflip(side), u = upflip(uplo);                                             \
            WITH_MAP(symm_row, PRE##NAME##_(&s, &u, &n, &m, al, (const R *)a, &lda,              \
                                            (const R *)b, &ldb, be, (R *)c, &ldc));              \
        }                                                                                        \
    }
DEFINE_CBLAS_CSYMM(z, symm, double, "ZSYMM ")
DEFINE_CBLAS_CSYMM(c, symm, float, "CSYMM ")
DEFINE_CBLAS_CSYMM(z, hemm, double, "ZHEMM ")
DEFINE_CBLAS_CSYMM(c, hemm, float, "CHEMM ")

#define DEFINE_CBLAS_CTRI(PRE, OP, R, NAME6)                                                     \
    void cblas_##PRE##OP(CBLAS_LAYOUT layout, CBLAS_SIDE side, CBLAS_UPLO uplo,                  \
                         CBLAS_TRANSPOSE ta, CBLAS_DIAG diag, int m, int n, const void *alpha,   \
                         const void *a, int lda, void *b, int ldb)                               \
    {                                                                                            \
        if (!lay_ok(layout)) { cbl_err(NAME6, 1); return; }                                      \
        if (!side_ok(side)) { cbl_err(NAME6, 2); return; }                                       \
        if (!up_ok(uplo)) { cbl_err(NAME6, 3); return; }                                         \
        if (!tr_ok(ta)) { cbl_err(NAME6, 4); return; }                                           \
        if (!diag_ok(diag)) { cbl_err(NAME6, 5); return; }                                       \
        char t = tr(ta), d = dg(diag);                                                           \
        const R *al = (const R *)alpha;                                                          \
        if (layout == CblasColMajor) {                                                           \
            char s = sd(side), u = up(uplo);                                                     \
            WITH_MAP(trsm_col, PRE##OP##_(&s, &u, &t, &d, &m, &n, al, (const R *)a, &lda,        \
                                          (R *)b, &ldb));                                        \
        } else {                                                                                 \
            char s = sdflip(side), u = upflip(uplo);                                             \
            WITH_MAP(trsm_row, PRE##OP##_(&s, &u, &t, &d, &n, &m, al, (const R *)a, &lda,        \
                                          (R *)b, &ldb));                                        \
        }                                                                                        \
    }
DEFINE_CBLAS_CTRI(z, trsm, double, "ZTRSM ")
DEFINE_CBLAS_CTRI(c, trsm, float, "CTRSM ")
DEFINE_CBLAS_CTRI(z, trmm, double, "ZTRMM ")
DEFINE_CBLAS_CTRI(c, trmm, float, "CTRMM ")

/* ---- GEMV: 1 layout 2 trans 3 m 4 n 5 alpha 6 A 7 lda 8 x 9 incx 10 beta 11 y 12 incy ---- */
static const int gemv_col[12] = {0, 2, 3, 4, 0, 0, 7, 0, 9, 0, 0, 12};
static const int gemv_row[12] = {0, 2, 4, 3, 0, 0, 7, 0, 9, 0, 0, 12};

void cblas_dgemv(CBLAS_LAYOUT layout, CBLAS_TRANSPOSE trans, int m, int n, double alpha,
                 const double *a, int lda, const double *x, int incx, double beta, double *y,
                 int incy)
{
    if (!lay_ok(layout)) { cbl_err("DGEMV ", 1); return; }
    if (!tr_ok(trans)) { cbl_err("DGEMV ", 2); return; }
    if (layout == CblasColMajor) {
        char t = tr(trans);
        WITH_MAP(gemv_col, dgemv_(&t, &m, &n, &alpha, a, &lda, x, &incx, &beta, y, &incy));
    } else {
        char t = trflip(trans);
        WITH_MAP(gemv_row, dgemv_(&t, &n, &m, &alpha, a, &lda, x, &incx, &beta, y, &incy));
    }
}

/* ---- GER: 1 layout 2 m 3 n 4 alpha 5 x 6 incx 7 y 8 incy 9 A 10 lda ---- */
static const int ger_col[10] = {0, 2, 3, 0, 0, 6, 0, 8, 0, 10};
static const int ger_row[10] = {0, 3, 2, 0, 0, 8, 0, 6, 0, 10};

void cblas_dger(CBLAS_LAYOUT layout, int m, int n, double alpha, const double *x, int incx,
                const double *y, int incy, double *a, int lda)
{
    if (!lay_ok(layout)) { cbl_err("DGER  ", 1); return; }
    if (layout == CblasColMajor)
        WITH_MAP(ger_col, dger_(&m, &n, &alpha, x, &incx, y, &incy, a, &lda));
    else
        WITH_MAP(ger_row, dger_(&n, &m, &alpha, y, &incy, x, &incx, a, &lda));
}

/* ---- TRSV: 1 layout 2 uplo 3 trans 4 diag 5 n 6 A 7 lda 8 x 9 incx ---- */
static const int trsv_map[9] = {0, 2, 3, 4, 5, 0, 7, 0, 9};

void cblas_dtrsv(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, CBLAS_DIAG diag,
                 int n, const double *a, int lda, double *x, int incx)
{
    if (!lay_ok(layout)) { cbl_err("DTRSV ", 1); return; }
    if (!up_ok(uplo)) { cbl_err("DTRSV ", 2); return; }
    if (!tr_ok(trans)) { cbl_err("DTRSV ", 3); return; }
    if (!diag_ok(diag)) { cbl_err("DTRSV ", 4); return; }
    char u = layout == CblasColMajor ? up(uplo) : upflip(uplo);
    char t = layout == CblasColMajor ? tr(trans) : trflip(trans);
    char d = dg(diag);
    WITH_MAP(trsv_map, dtrsv_(&u, &t, &d, &n, a, &lda, x, &incx));
}

/* ---- SYMV: 1 layout 2 uplo 3 n 4 alpha 5 A 6 lda 7 x 8 incx 9 beta 10 y 11 incy.
 * A row-major symmetric matrix is the column-major one of the other triangle. */
static const int symv_map[11] = {0, 2, 3, 0, 0, 6, 0, 8, 0, 0, 11};

void cblas_dsymv(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, int n, double alpha, const double *a,
                 int lda, const double *x, int incx, double beta, double *y, int incy)
{
    if (!lay_ok(layout)) { cbl_err("DSYMV ", 1); return; }
    if (!up_ok(uplo)) { cbl_err("DSYMV ", 2); return; }
    char u = layout == CblasColMajor ? up(uplo) : upflip(uplo);
    WITH_MAP(symv_map, dsymv_(&u, &n, &alpha, a, &lda, x, &incx, &beta, y, &incy));
}

/* ---- SYR: 1 layout 2 uplo 3 n 4 alpha 5 x 6 incx 7 A 8 lda ---- */
static const int syr_map[8] = {0, 2, 3, 0, 0, 6, 0, 8};

void cblas_dsyr(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, int n, double alpha, const double *x,
                int incx, double *a, int lda)
{
    if (!lay_ok(layout)) { cbl_err("DSYR  ", 1); return; }
    if (!up_ok(uplo)) { cbl_err("DSYR  ", 2); return; }
    char u = layout == CblasColMajor ? up(uplo) : upflip(uplo);
    WITH_MAP(syr_map, dsyr_(&u, &n, &alpha, x, &incx, a, &lda));
}

/* ---- SYR2: 1 layout 2 uplo 3 n 4 alpha 5 x 6 incx 7 y 8 incy 9 A 10 lda ---- */
static const int syr2_map[10] = {0, 2, 3, 0, 0, 6, 0, 8, 0, 10};

void cblas_dsyr2(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, int n, double alpha, const double *x,
                 int incx, const double *y, int incy, double *a, int lda)
{
    if (!lay_ok(layout)) { cbl_err("DSYR2 ", 1); return; }
    if (!up_ok(uplo)) { cbl_err("DSYR2 ", 2); return; }
    char u = layout == CblasColMajor ? up(uplo) : upflip(uplo);
    WITH_MAP(syr2_map, dsyr2_(&u, &n, &alpha, x, &incx, y, &incy, a, &lda));
}

/* ---- SPR: 1 layout 2 uplo 3 n 4 alpha 5 x 6 incx 7 Ap;
 *      SPR2: ... 7 y 8 incy 9 Ap.  A row-major packed triangle is the
 *      column-major packed other triangle. ---- */
static const int spr_map[6] = {0, 2, 3, 0, 0, 6};
static const int spr2_map[8] = {0, 2, 3, 0, 0, 6, 0, 8};

void cblas_dspr(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, int n, double alpha, const double *x,
                int incx, double *ap)
{
    if (!lay_ok(layout)) { cbl_err("DSPR  ", 1); return; }
    if (!up_ok(uplo)) { cbl_err("DSPR  ", 2); return; }
    char u = layout == CblasColMajor ? up(uplo) : upflip(uplo);
    WITH_MAP(spr_map, dspr_(&u, &n, &alpha, x, &incx, ap));
}

void cblas_dspr2(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, int n, double alpha, const double *x,
                 int incx, const double *y, int incy, double *ap)
{
    if (!lay_ok(layout)) { cbl_err("DSPR2 ", 1); return; }
    if (!up_ok(uplo)) { cbl_err("DSPR2 ", 2); return; }
    char u = layout == CblasColMajor ? up(uplo) : upflip(uplo);
    WITH_MAP(spr2_map, dspr2_(&u, &n, &alpha, x, &incx, y, &incy, ap));
}

/* ---- TRMV (positions as TRSV); TPMV / TPSV: 1 layout 2 uplo 3 trans 4 diag
 *      5 n 6 Ap 7 x 8 incx.  Row-major: the transposed system, other triangle. ---- */
static const int tpmv_map[8] = {0, 2, 3, 4, 5, 0, 0, 8};

void cblas_dtrmv(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, CBLAS_DIAG diag,
                 int n, const double *a, int lda, double *x, int incx)
{
    if (!lay_ok(layout)) { cbl_err("DTRMV ", 1); return; }
    if (!up_ok(uplo)) { cbl_err("DTRMV ", 2); return; }
    if (!tr_ok(trans)) { cbl_err("DTRMV ", 3); return; }
    if (!diag_ok(diag)) { cbl_err("DTRMV ", 4); return; }
    char u = layout == CblasColMajor ? up(uplo) : upflip(uplo);
    char t = layout == CblasColMajor ? tr(trans) : trflip(trans);
    char d = dg(diag);
    WITH_MAP(trsv_map, dtrmv_(&u, &t, &d, &n, a, &lda, x, &incx));
}

static void cblas_tp(int solve, CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans,
                     CBLAS_DIAG diag, int n, const double *ap, double *x, int incx)
{
    const char *nm = solve ? "DTPSV " : "DTPMV ";
    if (!lay_ok(layout)) { cbl_err(nm, 1); return; }
    if (!up_ok(uplo)) { cbl_err(nm, 2); return; }
    if (!tr_ok(trans)) { cbl_err(nm, 3); return; }
    if (!diag_ok(diag)) { cbl_err(nm, 4); return; }
    char u = layout == CblasColMajor ? up(uplo) : upflip(uplo);
    char t = layout == CblasColMajor ? tr(trans) : trflip(trans);
    char d = dg(diag);
    if (solve) WITH_MAP(tpmv_map, dtpsv_(&u, &t, &d, &n, ap, x, &incx));
    else WITH_MAP(tpmv_map, dtpmv_(&u, &t, &d, &n, ap, x, &incx));
}

void cblas_dtpmv(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, CBLAS_DIAG diag,
                 int n, const double *ap, double *x, int incx)
{
    cblas_tp(0, layout, uplo, trans, diag, n, ap, x, incx);
}

void cblas_dtpsv(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, CBLAS_DIAG diag,
                 int n, const double *ap, double *x, int incx)
{
    cblas_tp(1, layout, uplo, trans, diag, n, ap, x, incx);
}

/* ---- SPMV: 1 layout 2 uplo 3 n 4 alpha 5 Ap 6 x 7 incx 8 beta 9 y 10 incy ---- */
static const int spmv_map[10] = {0, 2, 3, 0, 0, 0, 7, 0, 0, 10};

void cblas_dspmv(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, int n, double alpha, const double *ap,
                 const double *x, int incx, double beta, double *y, int incy)
{
    if (!lay_ok(layout)) { cbl_err("DSPMV ", 1); return; }
    if (!up_ok(uplo)) { cbl_err("DSPMV ", 2); return; }
    char u = layout == CblasColMajor ? up(uplo) : upflip(uplo);
    WITH_MAP(spmv_map, dspmv_(&u, &n, &alpha, ap, x, &incx, &beta, y, &incy));
}

/* ---- GBMV: 1 layout 2 trans 3 M 4 N 5 KL 6 KU 7 alpha 8 A 9 lda 10 x 11 incx
 *      12 beta 13 y 14 incy.  Row-major band storage of A is the column-major
 *      band storage of A' (n x m, bands ku / kl). ---- */
static const int gbmv_col[14] = {0, 2, 3, 4, 5, 6, 0, 0, 9, 0, 11, 0, 0, 14};
static const int gbmv_row[14] = {0, 2, 4, 3, 6, 5, 0, 0, 9, 0, 11, 0, 0, 14};

void cblas_dgbmv(CBLAS_LAYOUT layout, CBLAS_TRANSPOSE trans, int m, int n, int kl, int ku,
                 double alpha, const double *a, int lda, const double *x, int incx, double beta,
                 double *y, int incy)
{
    if (!lay_ok(layout)) { cbl_err("DGBMV ", 1); return; }
    if (!tr_ok(trans)) { cbl_err("DGBMV ", 2); return; }
    if (layout == CblasColMajor) {
        char t = tr(trans);
        WITH_MAP(gbmv_col,
                 dgbmv_(&t, &m, &n, &kl, &ku, &alpha, a, &lda, x, &incx, &beta, y, &incy));
    } else {
        char t = trflip(trans);
        WITH_MAP(gbmv_row,
                 dgbmv_(&t, &n, &m, &ku, &kl, &alpha, a, &lda, x, &incx, &beta, y, &incy));
    }
}

/* ---- SBMV: 1 layout 2 uplo 3 n 4 k 5 alpha 6 A 7 lda 8 x 9 incx 10 beta 11 y
 *      12 incy.  Row-major: the other triangle of the same band bytes. ---- */
static const int sbmv_map[12] = {0, 2, 3, 4, 0, 0, 7, 0, 9, 0, 0, 12};

void cblas_dsbmv(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, int n, int k, double alpha, const double *a,
                 int lda, const double *x, int incx, double beta, double *y, int incy)
{
    if (!lay_ok(layout)) { cbl_err("DSBMV ", 1); return; }
    if (!up_ok(uplo)) { cbl_err("DSBMV ", 2); return; }
    char u = layout == CblasColMajor ? up(uplo) : upflip(uplo);
    WITH_MAP(sbmv_map, dsbmv_(&u, &n, &k, &alpha, a, &lda, x, &incx, &beta, y, &incy));
}

/* ---- TBMV / TBSV: 1 layout 2 uplo 3 trans 4 diag 5 n 6 k 7 A 8 lda 9 x
 *      10 incx.  Row-major: the transposed system, other triangle. ---- */
static const int tbmv_map[10] = {0, 2, 3, 4, 5, 6, 0, 8, 0, 10};

static void cblas_tb(int solve, CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans,
                     CBLAS_DIAG diag, int n, int k, const double *a, int lda, double *x, int incx)
{
    const char *nm = solve ? "DTBSV " : "DTBMV ";
    if (!lay_ok(layout)) { cbl_err(nm, 1); return; }
    if (!up_ok(uplo)) { cbl_err(nm, 2); return; }
    if (!tr_ok(trans)) { cbl_err(nm, 3); return; }
    if (!diag_ok(diag)) { cbl_err(nm, 4); return; }
    char u = layout == CblasColMajor ? up(uplo) : upflip(uplo);
    char t = layout == CblasColMajor ? tr(trans) : trflip(trans);
    char d = dg(diag);
    if (solve) WITH_MAP(tbmv_map, dtbsv_(&u, &t, &d, &n, &k, a, &lda, x, &incx));
    else WITH_MAP(tbmv_map, dtbmv_(&u, &t, &d, &n, &k, a, &lda, x, &incx));
}

void cblas_dtbmv(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, CBLAS_DIAG diag,
                 int n, int k, const double *a, int lda, double *x, int incx)
{
    cblas_tb(0, layout, uplo, trans, diag, n, k, a, lda, x, incx);
}

void cblas_dtbsv(CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans, CBLAS_DIAG diag,
                 int n, int k, const double *a, int lda, double *x, int incx)
{
    cblas_tb(1, layout, uplo, trans, diag, n, k, a, lda, x, incx);
}

/* ---- Level 1 (no layout argument) ---- */
void cblas_daxpy(int n, double alpha, const double *x, int incx, double *y, int incy)
{
    daxpy_(&n, &alpha, x, &incx, y, &incy);
}

double cblas_ddot(int n, const double *x, int incx, const double *y, int incy)
{
    return ddot_(&n, x, &incx, y, &incy);
}

double cblas_dnrm2(int n, const double *x, int incx) { return dnrm2_(&n, x, &incx); }

int cblas_idamax(int n, const double *x, int incx)
{
    int r = idamax_(&n, x, &incx);
    return r ? r - 1 : 0; /* CBLAS indices are 0-based */
}

int cblas_idamin(int n, const double *x, int incx)
{
    int r = idamin_(&n, x, &incx);
    return r ? r - 1 : 0; /* CBLAS indices are 0-based */
}

void cblas_drotm(int n, double *x, int incx, double *y, int incy, const double *param)
{
    drotm_(&n, x, &incx, y, &incy, param);
}

void cblas_dscal(int n, double alpha, double *x, int incx) { dscal_(&n, &alpha, x, &incx); }

void cblas_dcopy(int n, const double *x, int incx, double *y, int incy)
{
    dcopy_(&n, x, &incx, y, &incy);
}

void cblas_dswap(int n, double *x, int incx, double *y, int incy) { dswap_(&n, x, &incx, y, &incy); }

void cblas_drot(int n, double *x, int incx, double *y, int incy, double c, double s)
{
    drot_(&n, x, &incx, y, &incy, &c, &s);
}

double cblas_dasum(int n, const double *x, int incx) { return dasum_(&n, x, &incx); }
