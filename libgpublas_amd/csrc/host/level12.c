/*
 * level12.c -- Fortran-ABI Level-1 and Level-2 entry points on the hot
 * path: daxpy_, ddot_, dnrm2_, idamax_, dgemv_, dger_, dtrsv_, and dscal_,
 * dcopy_, dswap_, drot_, dasum_, dsymv_, dsyr_, dsyr2_, dspr_, dspr2_,
 * dtrmv_, dtpmv_, dtpsv_, dspmv_, dgbmv_, dsbmv_, dtbmv_, dtbsv_.
 *
 * The reference specifies these only through CBLAS wrappers it never builds
 * (blas_level1/axpy.cc:10-50, dot.cc:9-48, nrm2.cc:9-41, amax.cc:9-43,
 * blas_level2/gemv.cc:11-98, ger.cc:11-94, trsv.cc:11-88; meson.build:34-76).
 * Their behaviour is the netlib definition those wrappers forward to; the
 * argument checks are netlib's (INFO codes in the comments).  Scalar results
 * are written by the kernel into a pinned per-thread slot together with a
 * sequence number; the call polls for that number (a few microseconds less
 * than a stream synchronisation) and returns the value.
 *
 * Host fall-through: when none of a call's operands is a tracked
 * (device-resident) object and its length is below the GEMM cutoff, the call
 * goes to the host BLAS -- the reference's small-problem rule (gemm.cc:129-
 * 141) applied to vectors, where staging 2 x 4 KB through the host link
 * costs more than the arithmetic.  cutoff=0 sends everything to the GPU;
 * cutoff=auto decides by the measured crossover table of route.c.
 */
#include "internal.h"
#include <math.h>
#include <string.h>

#define MAXI(a, b) ((a) > (b) ? (a) : (b))

/* Level-1: the fixed rule compares the length with the cutoff; the crossover
 * table sees 2n flops over 24n (streaming: x, y in and out) or 16n (reduction:
 * x, y) bytes. */
static int l1_host(int kind, int n, const void *p, const void *q)
{
    const double nn = n;
    return gb_host_route(kind, gb_opt.cutoff > 0 && n < gb_opt.cutoff, 2 * nn,
                         (kind == GB_RK_STREAM ? 24 : 16) * nn, p, q, NULL);
}

/* ---- DAXPY: y := alpha*x + y ---- */
void daxpy_(const int *n, const double *alpha, const double *x, const int *incx, double *y,
            const int *incy)
{
    gb_enter();
    GB_COUNT(calls, 1);
    if (*n <= 0 || *alpha == 0.0) {
        gb_leave();
        return;
    }
    if (l1_host(GB_RK_STREAM, *n, x, y)) {
        void (*f)(const int *, const double *, const double *, const int *, double *,
                  const int *) = gb_host_fn("daxpy_");
        GB_COUNT(fallthrough, 1);
        f(n, alpha, x, incx, y, incy);
        gb_leave();
        return;
    }
    gb_require_device();
    void *s = gpublas_dev_stream();
    gb_operand ops[2];
    gb_acquire_vector(&ops[0], x, *n, *incx, 8, GB_IN, s);
    gb_acquire_vector(&ops[1], y, *n, *incy, 8, GB_INOUT, s);
    gb_check(gpublas_dev_daxpy(s, *n, *alpha, (const double *)ops[0].dev, *incx,
                               (double *)ops[1].dev, *incy),
             "daxpy launch");
    gb_finish(s, ops, 2);
    GB_COUNT(device_calls, 1);
    gb_leave();
}

/* DSCAL / DCOPY / DSWAP / DROT / DASUM: netlib definitions, which the
 * reference's CBLAS specs forward to (blas_level1/scal.cc:42, copy.cc:46,
 * swap.cc:12-29, rot.cc:50, asum.cc:48); same fall-through rule as DAXPY for small
 * untracked vectors. */
void dscal_(const int *n, const double *alpha, double *x, const int *incx)
{
    gb_enter();
    GB_COUNT(calls, 1);
    if (*n <= 0 || *incx <= 0) {
        gb_leave();
        return;
    }
    if (l1_host(GB_RK_STREAM, *n, x, NULL)) {
        void (*f)(const int *, const double *, double *, const int *) = gb_host_fn("dscal_");
        GB_COUNT(fallthrough, 1);
        f(n, alpha, x, incx);
        gb_leave();
        return;
    }
    gb_require_device();
    void *s = gpublas_dev_stream();
    gb_operand ops[1];
    gb_acquire_vector(&ops[0], x, *n, *incx, 8, GB_INOUT, s);
    gb_check(gpublas_dev_dscal(s, *n, *alpha, (double *)ops[0].dev, *incx), "dscal launch");
    gb_finish(s, ops, 1);
    GB_COUNT(device_calls, 1);
    gb_leave();
}

void dcopy_(const int *n, const double *x, const int *incx, double *y, const int *incy)
{
    gb_enter();
    GB_COUNT(calls, 1);
    if (*n <= 0) {
        gb_leave();
        return;
    }
    if (l1_host(GB_RK_STREAM, *n, x, y)) {
        void (*f)(const int *, const double *, const int *, double *, const int *) =
            gb_host_fn("dcopy_");
        GB_COUNT(fallthrough, 1);
        f(n, x, incx, y, incy);
        gb_leave();
        return;
    }
    gb_require_device();
    void *s = gpublas_dev_stream();
    gb_operand ops[2];
    gb_acquire_vector(&ops[0], x, *n, *incx, 8, GB_IN, s);
    gb_acquire_vector(&ops[1], y, *n, *incy, 8, GB_OUT, s);
    gb_check(gpublas_dev_dcopy(s, *n, (const double *)ops[0].dev, *incx, (double *)ops[1].dev,
                               *incy),
             "dcopy launch");
    gb_finish(s, ops, 2);
    GB_COUNT(device_calls, 1);
    gb_leave();
}

void dswap_(const int *n, double *x, const int *incx, double *y, const int *incy)
{
    gb_enter();
    GB_COUNT(calls, 1);
    if (*n <= 0) {
        gb_leave();
        return;
    }
    if (l1_host(GB_RK_STREAM, *n, x, y)) {
        void (*f)(const int *, double *, const int *, double *, const int *) = gb_host_fn("dswap_");
        GB_COUNT(fallthrough, 1);
        f(n, x, incx, y, incy);
        gb_leave();
        return;
    }
    gb_require_device();
    void *s = gpublas_dev_stream();
    gb_operand ops[2];
    gb_acquire_vector(&ops[0], x, *n, *incx, 8, GB_INOUT, s);
    gb_acquire_vector(&ops[1], y, *n, *incy, 8, GB_INOUT, s);
    gb_check(gpublas_dev_dswap(s, *n, (double *)ops[0].dev, *incx, (double *)ops[1].dev, *incy),
             "dswap launch");
    gb_finish(s, ops, 2);
    GB_COUNT(device_calls, 1);
    gb_leave();
}

void drot_(const int *n, double *x, const int *incx, double *y, const int *incy, const double *c,
           const double *sn)
{
    gb_enter();
    GB_COUNT(calls, 1);
    if (*n <= 0) {
        gb_leave();
        return;
    }
    if (l1_host(GB_RK_STREAM, *n, x, y)) {
        void (*f)(const int *, double *, const int *, double *, const int *, const double *,
                  const double *) = gb_host_fn("drot_");
        GB_COUNT(fallthrough, 1);
        f(n, x, incx, y, incy, c, sn);
        gb_leave();
        return;
    }
    gb_require_device();
    void *s = gpublas_dev_stream();
    gb_operand ops[2];
    gb_acquire_vector(&ops[0], x, *n, *incx, 8, GB_INOUT, s);
    gb_acquire_vector(&ops[1], y, *n, *incy, 8, GB_INOUT, s);
    gb_check(gpublas_dev_drot(s, *n, (double *)ops[0].dev, *incx, (double *)ops[1].dev, *incy, *c,
                              *sn),
             "drot launch");
    gb_finish(s, ops, 2);
    GB_COUNT(device_calls, 1);
    gb_leave();
}

/* Option `resident`: a short reduction whose operands are all tracked objects
 * goes to the resident server (no kernel launch).  Returns 1 when served; 0
 * leaves the call to the launched path (operands stay acquired). */
static int served_reduce(void *s, int op, int n, gb_operand *ops, int nops, int incx, int incy,
                         double *dr, int *ir)
{
    if (!gb_opt.resident) return 0;
    /* the server reads the operands at once: their writers (and uploads) must
     * be known to have ended; otherwise a launch queued behind them is as quick */
    if (!gb_writers_done(ops, nops)) return 0;
    if (!gpublas_dev_resident_active() && gpublas_dev_resident_start() != 0) return 0;
    if (gpublas_dev_resident_reduce(s, op, n, (const double *)ops[0].dev, incx,
                                    nops > 1 ? (const double *)ops[1].dev : NULL, incy, 1, dr,
                                    ir) != 0)
        return 0;
    gb_finish_served(s, ops, nops);
    GB_COUNT(device_calls, 1);
    GB_COUNT(resident, 1);
    return 1;
}

double dasum_(const int *n, const double *x, const int *incx)
{
    gb_enter();
    GB_COUNT(calls, 1);
    if (*n <= 0 || *incx <= 0) {
        gb_leave();
        return 0.0;
    }
    if (l1_host(GB_RK_REDUCE, *n, x, NULL)) {
        double (*f)(const int *, const double *, const int *) = gb_host_fn("dasum_");
        GB_COUNT(fallthrough, 1);
        double r = f(n, x, incx);
        gb_leave();
        return r;
    }
    gb_require_device();
    void *s = gpublas_dev_stream();
    gb_operand ops[1];
    gb_acquire_vector(&ops[0], x, *n, *incx, 8, GB_IN, s);
    double served;
    if (served_reduce(s, 4, *n, ops, 1, *incx, 0, &served, NULL)) {
        gb_leave();
        return served;
    }
    double *slot = (double *)gpublas_dev_result_slot();
    const uint64_t seq = gpublas_dev_scalar_next();
    gb_check(gpublas_dev_dasum(s, *n, (const double *)ops[0].dev, *incx, slot), "dasum launch");
    gb_finish_scalar(s, ops, 1, seq);
    double r = *(volatile double *)slot;
    GB_COUNT(device_calls, 1);
    gb_leave();
    return r;
}

double ddot_(const int *n, const double *x, const int *incx, const double *y, const int *incy)
{
    gb_enter();
    GB_COUNT(calls, 1);
    if (*n <= 0) {
        gb_leave();
        return 0.0;
    }
    if (l1_host(GB_RK_REDUCE, *n, x, y)) {
        double (*f)(const int *, const double *, const int *, const double *, const int *) =
            gb_host_fn("ddot_");
        GB_COUNT(fallthrough, 1);
        double r = f(n, x, incx, y, incy);
        gb_leave();
        return r;
    }
    gb_require_device();
    void *s = gpublas_dev_stream();
    gb_operand ops[2];
    gb_acquire_vector(&ops[0], x, *n, *incx, 8, GB_IN, s);
    gb_acquire_vector(&ops[1], y, *n, *incy, 8, GB_IN, s);
    double served;
    if (served_reduce(s, 0, *n, ops, 2, *incx, *incy, &served, NULL)) {
        gb_leave();
        return served;
    }
    double *slot = (double *)gpublas_dev_result_slot();
    const uint64_t seq = gpublas_dev_scalar_next();
    gb_check(gpublas_dev_ddot(s, *n, (const double *)ops[0].dev, *incx,
                              (const double *)ops[1].dev, *incy, slot),
             "ddot launch");
    gb_finish_scalar(s, ops, 2, seq);
    double r = *(volatile double *)slot;
    GB_COUNT(device_calls, 1);
    gb_leave();
    return r;
}

/* ---- DNRM2 (n <= 0 -> 0; incx == 0 repeats x(1); incx < 0 same set) ---- */
double dnrm2_(const int *n, const double *x, const int *incx)
{
    gb_enter();
    GB_COUNT(calls, 1);
    if (*n <= 0) {
        gb_leave();
        return 0.0;
    }
    if (l1_host(GB_RK_REDUCE, *n, x, NULL)) {
        double (*f)(const int *, const double *, const int *) = gb_host_fn("dnrm2_");
        GB_COUNT(fallthrough, 1);
        double r = f(n, x, incx);
        gb_leave();
        return r;
    }
    gb_require_device();
    void *s = gpublas_dev_stream();
    gb_operand op;
    gb_acquire_vector(&op, x, *n, *incx, 8, GB_IN, s);
    double served;
    if (served_reduce(s, 1, *n, &op, 1, *incx, 0, &served, NULL)) {
        gb_leave();
        return served;
    }
    double *slot = (double *)gpublas_dev_result_slot();
    const uint64_t seq = gpublas_dev_scalar_next();
    gb_check(gpublas_dev_dnrm2(s, *n, (const double *)op.dev, *incx, slot), "dnrm2 launch");
    gb_finish_scalar(s, &op, 1, seq);
    double r = *(volatile double *)slot;
    GB_COUNT(device_calls, 1);
    gb_leave();
    return r;
}

/* ---- IDAMAX (1-based, first maximal |x_i|; n < 1 or incx <= 0 -> 0) ---- */
int idamax_(const int *n, const double *x, const int *incx)
{
    gb_enter();
    GB_COUNT(calls, 1);
    if (*n < 1 || *incx <= 0) {
        gb_leave();
        return 0;
    }
    if (*n == 1) {
        gb_leave();
        return 1;
    }
    if (l1_host(GB_RK_REDUCE, *n, x, NULL)) {
        int (*f)(const int *, const double *, const int *) = gb_host_fn("idamax_");
        GB_COUNT(fallthrough, 1);
        int r = f(n, x, incx);
        gb_leave();
        return r;
    }
    gb_require_device();
    void *s = gpublas_dev_stream();
    gb_operand op;
    gb_acquire_vector(&op, x, *n, *incx, 8, GB_IN, s);
    int served;
    if (served_reduce(s, 2, *n, &op, 1, *incx, 0, NULL, &served)) {
        gb_leave();
        return served;
    }
    int *slot = (int *)gpublas_dev_result_slot();
    const uint64_t seq = gpublas_dev_scalar_next();
    gb_check(gpublas_dev_idamax(s, *n, (const double *)op.dev, *incx, slot), "idamax launch");
    gb_finish_scalar(s, &op, 1, seq);
    int r = *(volatile int *)slot;
    GB_COUNT(device_calls, 1);
    gb_leave();
    return r;
}

/* IDAMIN: the extension the reference's cblas_idamin forwards to
 * (blas_level1/amin.cc:40, cublasIdamin); the IDAMAX rules with the minimum. */
int idamin_(const int *n, const double *x, const int *incx)
{
    gb_enter();
    GB_COUNT(calls, 1);
    if (*n < 1 || *incx <= 0) {
        gb_leave();
        return 0;
    }
    if (*n == 1) {
        gb_leave();
        return 1;
    }
    if (l1_host(GB_RK_REDUCE, *n, x, NULL)) {
        int (*f)(const int *, const double *, const int *) = gb_host_fn("idamin_");
        GB_COUNT(fallthrough, 1);
        int r = f(n, x, incx);
        gb_leave();
        return r;
    }
    gb_require_device();
    void *s = gpublas_dev_stream();
    gb_operand op;
    gb_acquire_vector(&op, x, *n, *incx, 8, GB_IN, s);
    int served;
    if (served_reduce(s, 3, *n, &op, 1, *incx, 0, NULL, &served)) {
        gb_leave();
        return served;
    }
    int *slot = (int *)gpublas_dev_result_slot();
    const uint64_t seq = gpublas_dev_scalar_next();
    gb_check(gpublas_dev_idamin(s, *n, (const double *)op.dev, *incx, slot), "idamin launch");
    gb_finish_scalar(s, &op, 1, seq);
    int r = *(volatile int *)slot;
    GB_COUNT(device_calls, 1);
    gb_leave();
    return r;
}

/* DROTM: netlib DROTM, the definition blas_level1/rotm.cc:51 forwards to;
 * param = (flag, h11, h21, h12, h22) stays on the host. */
void drotm_(const int *n, double *x, const int *incx, double *y, const int *incy,
            const double *param)
{
    gb_enter();
    GB_COUNT(calls, 1);
    if (*n <= 0 || param[0] == -2.0) {
        gb_leave();
        return;
    }
    if (l1_host(GB_RK_STREAM, *n, x, y)) {
        void (*f)(const int *, double *, const int *, double *, const int *, const double *) =
            gb_host_fn("drotm_");
        GB_COUNT(fallthrough, 1);
        f(n, x, incx, y, incy, param);
        gb_leave();
        return;
    }
    gb_require_device();
    void *s = gpublas_dev_stream();
    gb_operand ops[2];
    gb_acquire_vector(&ops[0], x, *n, *incx, 8, GB_INOUT, s);
    gb_acquire_vector(&ops[1], y, *n, *incy, 8, GB_INOUT, s);
    gb_check(gpublas_dev_drotm(s, *n, (double *)ops[0].dev, *incx, (double *)ops[1].dev, *incy,
                               param),
             "drotm launch");
    gb_finish(s, ops, 2);
    GB_COUNT(device_calls, 1);
    gb_leave();
}

/* ---- DGEMV: y := alpha*op(A)*x + beta*y
 * INFO: 1 trans, 2 m, 3 n, 6 lda, 8 incx, 11 incy ---- */
void dgemv_(const char *trans, const int *m, const int *n, const double *alpha, const double *a,
            const int *lda, const double *x, const int *incx, const double *beta, double *y,
            const int *incy)
{
    gb_enter();
    GB_COUNT(calls, 1);
    const int notr = gb_lsame(*trans, 'N');
    int info = 0;
    if (!notr && !gb_lsame(*trans, 'T') && !gb_lsame(*trans, 'C')) info = 1;
    else if (*m < 0) info = 2;
    else if (*n < 0) info = 3;
    else if (*lda < MAXI(1, *m)) info = 6;
    else if (*incx == 0) info = 8;
    else if (*incy == 0) info = 11;
    if (info) {
        gb_xerbla("DGEMV ", info);
        gb_leave();
        return;
    }
    if (*m == 0 || *n == 0 || (*alpha == 0.0 && *beta == 1.0)) {
        gb_leave();
        return;
    }
    const int lenx = notr ? *n : *m, leny = notr ? *m : *n;
    if (gb_host_route(GB_RK_DGEMV, gb_opt.cutoff > 0 && MAXI(*m, *n) < gb_opt.cutoff,
                      2.0 * *m * *n, 8.0 * ((double)*m * *n + *m + *n), a, x, y)) {
        void (*f)(const char *, const int *, const int *, const double *, const double *,
                  const int *, const double *, const int *, const double *, double *,
                  const int *) = gb_host_fn("dgemv_");
        GB_COUNT(fallthrough, 1);
        f(trans, m, n, alpha, a, lda, x, incx, beta, y, incy);
        gb_leave();
        return;
    }
    if (*alpha != 0.0 && gb_opt.serve &&
        gb_serve_dgemv(*trans, *m, *n, *alpha, a, *lda, x, *incx, *beta, y, *incy)) {
        GB_COUNT(device_calls, 1);
        gb_leave();
        return;
    }
    gb_require_device();
    void *s = gpublas_dev_stream();
    gb_operand ops[3];
    memset(ops, 0, sizeof ops);
    if (*alpha != 0.0) {
        gb_acquire_matrix(&ops[0], a, *lda, *m, *n, 8, GB_IN, s);
        gb_acquire_vector(&ops[1], x, lenx, *incx, 8, GB_IN, s);
    }
    gb_acquire_vector(&ops[2], y, leny, *incy, 8, *beta == 0.0 ? GB_OUT : GB_INOUT, s);
    gb_check(gpublas_dev_dgemv(s, *trans, *m, *n, *alpha, (const double *)ops[0].dev, *lda,
                               (const double *)ops[1].dev, *incx, *beta, (double *)ops[2].dev,
                               *incy),
             "dgemv launch");
    gb_finish(s, ops, 3);
    GB_COUNT(device_calls, 1);
    gb_leave();
}

/* ---- DGER: A := alpha*x*y**T + A
 * INFO: 1 m, 2 n, 5 incx, 7 incy, 9 lda ---- */
void dger_(const int *m, const int *n, const double *alpha, const double *x, const int *incx,
           const double *y, const int *incy, double *a, const int *lda)
{
    gb_enter();
    GB_COUNT(calls, 1);
    int info = 0;
    if (*m < 0) info = 1;
    else if (*n < 0) info = 2;
    else if (*incx == 0) info = 5;
    else if (*incy == 0) info = 7;
    else if (*lda < MAXI(1, *m)) info = 9;
    if (info) {
        gb_xerbla("DGER  ", info);
        gb_leave();
        return;
    }
    if (*m == 0 || *n == 0 || *alpha == 0.0) {
        gb_leave();
        return;
    }
    if (gb_host_route(GB_RK_DGER, gb_opt.cutoff > 0 && MAXI(*m, *n) < gb_opt.cutoff,
                      2.0 * *m * *n, 8.0 * (2.0 * *m * *n + *m + *n), a, x, y)) {
        void (*f)(const int *, const int *, const double *, const double *, const int *,
                  const double *, const int *, double *, const int *) = gb_host_fn("dger_");
        GB_COUNT(fallthrough, 1);
        f(m, n, alpha, x, incx, y, incy, a, lda);
        gb_leave();
        return;
    }
    gb_require_device();
    void *s = gpublas_dev_stream();
    gb_operand ops[3];
    gb_acquire_vector(&ops[0], x, *m, *incx, 8, GB_IN, s);
    gb_acquire_vector(&ops[1], y, *n, *incy, 8, GB_IN, s);
    gb_acquire_matrix(&ops[2], a, *lda, *m, *n, 8, GB_INOUT, s);
    gb_check(gpublas_dev_dger(s, *m, *n, *alpha, (const double *)ops[0].dev, *incx,
                              (const double *)ops[1].dev, *incy, (double *)ops[2].dev, *lda),
             "dger launch");
    gb_finish(s, ops, 3);
    GB_COUNT(device_calls, 1);
    gb_leave();
}

/* ---- DTRSV: solve op(A)*x = b in place
 * INFO: 1 uplo, 2 trans, 3 diag, 4 n, 6 lda, 8 incx ---- */
void dtrsv_(const char *uplo, const char *trans, const char *diag, const int *n, const double *a,
            const int *lda, double *x, const int *incx)
{
    gb_enter();
    GB_COUNT(calls, 1);
    int info = 0;
    if (!gb_lsame(*uplo, 'U') && !gb_lsame(*uplo, 'L')) info = 1;
    else if (!gb_lsame(*trans, 'N') && !gb_lsame(*trans, 'T') && !gb_lsame(*trans, 'C')) info = 2;
    else if (!gb_lsame(*diag, 'U') && !gb_lsame(*diag, 'N')) info = 3;
    else if (*n < 0) info = 4;
    else if (*lda < MAXI(1, *n)) info = 6;
    else if (*incx == 0) info = 8;
    if (info) {
        gb_xerbla("DTRSV ", info);
        gb_leave();
        return;
    }
    if (*n == 0) {
        gb_leave();
        return;
    }
    if (gb_host_route(GB_RK_DTRSV, gb_opt.cutoff > 0 && *n < gb_opt.cutoff, (double)*n * *n,
                      8.0 * ((double)*n * *n + 2.0 * *n), a, x, NULL)) {
        void (*f)(const char *, const char *, const char *, const int *, const double *,
                  const int *, double *, const int *) = gb_host_fn("dtrsv_");
        GB_COUNT(fallthrough, 1);
        f(uplo, trans, diag, n, a, lda, x, incx);
        gb_leave();
        return;
    }
    gb_require_device();
    void *s = gpublas_dev_stream();
    gb_operand ops[2];
    gb_acquire_matrix(&ops[0], a, *lda, *n, *n, 8, GB_IN, s);
    gb_acquire_vector(&ops[1], x, *n, *incx, 8, GB_INOUT, s);
    gb_check(gpublas_dev_dtrsv(s, *uplo, *trans, *diag, *n, (const double *)ops[0].dev, *lda,
                               (double *)ops[1].dev, *incx),
             "dtrsv launch");
    gb_finish(s, ops, 2);
    GB_COUNT(device_calls, 1);
    gb_leave();
}

/* ---- DSYMV: y := alpha*A*x + beta*y, A symmetric (uplo triangle)
 * netlib DSYMV, the definition blas_level2/symv.cc:11-75 forwards to.
 * INFO: 1 uplo, 2 n, 5 lda, 7 incx, 10 incy ---- */
void dsymv_(const char *uplo, const int *n, const double *alpha, const double *a, const int *lda,
            const double *x, const int *incx, const double *beta, double *y, const int *incy)
{
    gb_enter();
    GB_COUNT(calls, 1);
    int info = 0;
    if (!gb_lsame(*uplo, 'U') && !gb_lsame(*uplo, 'L')) info = 1;
    else if (*n < 0) info = 2;
    else if (*lda < MAXI(1, *n)) info = 5;
    else if (*incx == 0) info = 7;
    else if (*incy == 0) info = 10;
    if (info) {
        gb_xerbla("DSYMV ", info);
        gb_leave();
        return;
    }
    if (*n == 0 || (*alpha == 0.0 && *beta == 1.0)) {
        gb_leave();
        return;
    }
    const double nn = *n;
    if (gb_host_route(GB_RK_DSYMV, gb_opt.cutoff > 0 && *n < gb_opt.cutoff, 2 * nn * nn,
                      8.0 * (nn * nn + 3 * nn), a, x, y)) {
        void (*f)(const char *, const int *, const double *, const double *, const int *,
                  const double *, const int *, const double *, double *, const int *) =
            gb_host_fn("dsymv_");
        GB_COUNT(fallthrough, 1);
        f(uplo, n, alpha, a, lda, x, incx, beta, y, incy);
        gb_leave();
        return;
    }
    gb_require_device();
    void *s = gpublas_dev_stream();
    gb_operand ops[3];
    memset(ops, 0, sizeof ops);
    if (*alpha != 0.0) {
        gb_acquire_matrix(&ops[0], a, *lda, *n, *n, 8, GB_IN, s);
        gb_acquire_vector(&ops[1], x, *n, *incx, 8, GB_IN, s);
    }
    gb_acquire_vector(&ops[2], y, *n, *incy, 8, *beta == 0.0 ? GB_OUT : GB_INOUT, s);
    gb_check(gpublas_dev_dsymv(s, *uplo, *n, *alpha, (const double *)ops[0].dev, *lda,
                               (const double *)ops[1].dev, *incx, *beta, (double *)ops[2].dev,
                               *incy),
             "dsymv launch");
    gb_finish(s, ops, 3);
    GB_COUNT(device_calls, 1);
    gb_leave();
}

/* ---- DSYR: A := alpha*x*x**T + A on the uplo triangle
 * netlib DSYR, the definition blas_level2/syr.cc:11-68 forwards to.
 * INFO: 1 uplo, 2 n, 5 incx, 7 lda ---- */
void dsyr_(const char *uplo, const int *n, const double *alpha, const double *x, const int *incx,
           double *a, const int *lda)
{
    gb_enter();
    GB_COUNT(calls, 1);
    int info = 0;
    if (!gb_lsame(*uplo, 'U') && !gb_lsame(*uplo, 'L')) info = 1;
    else if (*n < 0) info = 2;
    else if (*incx == 0) info = 5;
    else if (*lda < MAXI(1, *n)) info = 7;
    if (info) {
        gb_xerbla("DSYR  ", info);
        gb_leave();
        return;
    }
    if (*n == 0 || *alpha == 0.0) {
        gb_leave();
        return;
    }
    const double nn = *n;
    if (gb_host_route(GB_RK_DSYR, gb_opt.cutoff > 0 && *n < gb_opt.cutoff, nn * nn,
                      8.0 * (2 * nn * nn + nn), a, x, NULL)) {
        void (*f)(const char *, const int *, const double *, const double *, const int *,
                  double *, const int *) = gb_host_fn("dsyr_");
        GB_COUNT(fallthrough, 1);
        f(uplo, n, alpha, x, incx, a, lda);
        gb_leave();
        return;
    }
    gb_require_device();
    void *s = gpublas_dev_stream();
    gb_operand ops[2];
    memset(ops, 0, sizeof ops);
    gb_acquire_vector(&ops[0], x, *n, *incx, 8, GB_IN, s);
    /* the other triangle round-trips unchanged */
    gb_acquire_matrix(&ops[1], a, *lda, *n, *n, 8, GB_INOUT, s);
    gb_check(gpublas_dev_dsyr(s, *uplo, *n, *alpha, (const double *)ops[0].dev, *incx,
                              (double *)ops[1].dev, *lda),
             "dsyr launch");
    gb_finish(s, ops, 2);
    GB_COUNT(device_calls, 1);
    gb_leave();
}

/* ---- DSYR2: A := alpha*x*y**T + alpha*y*x**T + A on the uplo triangle
 * netlib DSYR2 (the definition blas_level2/syr2.cc forwards to).
 * INFO: 1 uplo, 2 n, 5 incx, 7 incy, 9 lda ---- */
void dsyr2_(const char *uplo, const int *n, const double *alpha, const double *x,
            const int *incx, const double *y, const int *incy, double *a, const int *lda)
{
    gb_enter();
    GB_COUNT(calls, 1);
    int info = 0;
    if (!gb_lsame(*uplo, 'U') && !gb_lsame(*uplo, 'L')) info = 1;
    else if (*n < 0) info = 2;
    else if (*incx == 0) info = 5;
    else if (*incy == 0) info = 7;
    else if (*lda < MAXI(1, *n)) info = 9;
    if (info) {
        gb_xerbla("DSYR2 ", info);
        gb_leave();
        return;
    }
    if (*n == 0 || *alpha == 0.0) {
        gb_leave();
        return;
    }
    const double nn = *n;
    if (gb_host_route(GB_RK_DSYR, gb_opt.cutoff > 0 && *n < gb_opt.cutoff, 2 * nn * nn,
                      8.0 * (2 * nn * nn + 2 * nn), a, x, y)) {
        void (*f)(const char *, const int *, const double *, const double *, const int *,
                  const double *, const int *, double *, const int *) = gb_host_fn("dsyr2_");
        GB_COUNT(fallthrough, 1);
        f(uplo, n, alpha, x, incx, y, incy, a, lda);
        gb_leave();
        return;
    }
    gb_require_device();
    void *s = gpublas_dev_stream();
    gb_operand ops[3];
    memset(ops, 0, sizeof ops);
    gb_acquire_vector(&ops[0], x, *n, *incx, 8, GB_IN, s);
    gb_acquire_vector(&ops[1], y, *n, *incy, 8, GB_IN, s);
    gb_acquire_matrix(&ops[2], a, *lda, *n, *n, 8, GB_INOUT, s);
    gb_check(gpublas_dev_dsyr2(s, *uplo, *n, *alpha, (const double *)ops[0].dev, *incx,
                               (const double *)ops[1].dev, *incy, (double *)ops[2].dev, *lda),
             "dsyr2 launch");
    gb_finish(s, ops, 3);
    GB_COUNT(device_calls, 1);
    gb_leave();
}

/* ---- DSPR / DSPR2: packed rank-1 / rank-2 updates (netlib DSPR, DSPR2; the
 * definitions blas_level2/spr.cc and spr2.cc forward to).  The packed
 * triangle is n(n+1)/2 contiguous elements.
 * INFO: 1 uplo, 2 n, 5 incx, 7 incy (DSPR2) ---- */
static void spr_common(int two, const char *uplo, const int *n, const double *alpha,
                       const double *x, const int *incx, const double *y, const int *incy,
                       double *ap)
{
    gb_enter();
    GB_COUNT(calls, 1);
    int info = 0;
    if (!gb_lsame(*uplo, 'U') && !gb_lsame(*uplo, 'L')) info = 1;
    else if (*n < 0) info = 2;
    else if (*incx == 0) info = 5;
    else if (two && *incy == 0) info = 7;
    if (info) {
        gb_xerbla(two ? "DSPR2 " : "DSPR  ", info);
        gb_leave();
        return;
    }
    if (*n == 0 || *alpha == 0.0) {
        gb_leave();
        return;
    }
    const double nn = *n, len = nn * (nn + 1) / 2;
    if (gb_host_route(GB_RK_DSYR, gb_opt.cutoff > 0 && *n < gb_opt.cutoff,
                      (two ? 2 : 1) * nn * nn, 8.0 * (2 * len + (two ? 2 : 1) * nn), ap, x,
                      two ? y : NULL)) {
        GB_COUNT(fallthrough, 1);
        if (two) {
            void (*f)(const char *, const int *, const double *, const double *, const int *,
                      const double *, const int *, double *) = gb_host_fn("dspr2_");
            f(uplo, n, alpha, x, incx, y, incy, ap);
        } else {
            void (*f)(const char *, const int *, const double *, const double *, const int *,
                      double *) = gb_host_fn("dspr_");
            f(uplo, n, alpha, x, incx, ap);
        }
        gb_leave();
        return;
    }
    gb_require_device();
    void *s = gpublas_dev_stream();
    gb_operand ops[3];
    memset(ops, 0, sizeof ops);
    gb_acquire_vector(&ops[0], x, *n, *incx, 8, GB_IN, s);
    if (two) gb_acquire_vector(&ops[1], y, *n, *incy, 8, GB_IN, s);
    gb_acquire_packed(&ops[2], ap, *n, 8, GB_INOUT, s);
    int e = two ? gpublas_dev_dspr2(s, *uplo, *n, *alpha, (const double *)ops[0].dev, *incx,
                                    (const double *)ops[1].dev, *incy, (double *)ops[2].dev)
                : gpublas_dev_dspr(s, *uplo, *n, *alpha, (const double *)ops[0].dev, *incx,
                                   (double *)ops[2].dev);
    gb_check(e, two ? "dspr2 launch" : "dspr launch");
    gb_finish(s, ops, 3);
    GB_COUNT(device_calls, 1);
    gb_leave();
}

void dspr_(const char *uplo, const int *n, const double *alpha, const double *x, const int *incx,
           double *ap)
{
    spr_common(0, uplo, n, alpha, x, incx, NULL, NULL, ap);
}

void dspr2_(const char *uplo, const int *n, const double *alpha, const double *x,
            const int *incx, const double *y, const int *incy, double *ap)
{
    spr_common(1, uplo, n, alpha, x, incx, y, incy, ap);
}

/* ---- DTRMV / DTPMV / DTPSV: x := op(A)*x and the packed solve (netlib
 * DTRMV, DTPMV, DTPSV; the definitions blas_level2/trmv.cc, tpmv.cc, tpsv.cc
 * forward to).  INFO: 1 uplo, 2 trans, 3 diag, 4 n, 6 lda (DTRMV), 8 incx
 * (DTRMV) / 7 incx (packed) ---- */
enum { TRI_MV = 0, TRI_PMV = 1, TRI_PSV = 2 };
static void tri_common(int kind, const char *uplo, const char *trans, const char *diag,
                       const int *n, const double *a, const int *lda, double *x, const int *incx)
{
    static const char *names[3] = {"DTRMV ", "DTPMV ", "DTPSV "};
    gb_enter();
    GB_COUNT(calls, 1);
    int info = 0;
    if (!gb_lsame(*uplo, 'U') && !gb_lsame(*uplo, 'L')) info = 1;
    else if (!gb_lsame(*trans, 'N') && !gb_lsame(*trans, 'T') && !gb_lsame(*trans, 'C')) info = 2;
    else if (!gb_lsame(*diag, 'U') && !gb_lsame(*diag, 'N')) info = 3;
    else if (*n < 0) info = 4;
    else if (kind == TRI_MV && *lda < MAXI(1, *n)) info = 6;
    else if (*incx == 0) info = kind == TRI_MV ? 8 : 7;
    if (info) {
        gb_xerbla(names[kind], info);
        gb_leave();
        return;
    }
    if (*n == 0) {
        gb_leave();
        return;
    }
    const double nn = *n, tri = nn * (nn + 1) / 2;
    if (gb_host_route(GB_RK_DTRSV, gb_opt.cutoff > 0 && *n < gb_opt.cutoff, nn * nn,
                      8.0 * (tri + 2 * nn), a, x, NULL)) {
        GB_COUNT(fallthrough, 1);
        if (kind == TRI_MV) {
            void (*f)(const char *, const char *, const char *, const int *, const double *,
                      const int *, double *, const int *) = gb_host_fn("dtrmv_");
            f(uplo, trans, diag, n, a, lda, x, incx);
        } else {
            void (*f)(const char *, const char *, const char *, const int *, const double *,
                      double *, const int *) = gb_host_fn(kind == TRI_PMV ? "dtpmv_" : "dtpsv_");
            f(uplo, trans, diag, n, a, x, incx);
        }
        gb_leave();
        return;
    }
    gb_require_device();
    void *s = gpublas_dev_stream();
    gb_operand ops[2];
    memset(ops, 0, sizeof ops);
    if (kind == TRI_MV) gb_acquire_matrix(&ops[0], a, *lda, *n, *n, 8, GB_IN, s);
    else gb_acquire_packed(&ops[0], a, *n, 8, GB_IN, s);
    gb_acquire_vector(&ops[1], x, *n, *incx, 8, GB_INOUT, s);
    int e;
    if (kind == TRI_MV)
        e = gpublas_dev_dtrmv(s, *uplo, *trans, *diag, *n, (const double *)ops[0].dev, *lda,
                              (double *)ops[1].dev, *incx);
    else if (kind == TRI_PMV)
        e = gpublas_dev_dtpmv(s, *uplo, *trans, *diag, *n, (const double *)ops[0].dev,
                              (double *)ops[1].dev, *incx);
    else
        e = gpublas_dev_dtpsv(s, *uplo, *trans, *diag, *n, (const double *)ops[0].dev,
                              (double *)ops[1].dev, *incx);
    gb_check(e, "triangular level-2 launch");
    gb_finish(s, ops, 2);
    GB_COUNT(device_calls, 1);
    gb_leave();
}

void dtrmv_(const char *uplo, const char *trans, const char *diag, const int *n, const double *a,
            const int *lda, double *x, const int *incx)
{
    tri_common(TRI_MV, uplo, trans, diag, n, a, lda, x, incx);
}

void dtpmv_(const char *uplo, const char *trans, const char *diag, const int *n, const double *ap,
            double *x, const int *incx)
{
    tri_common(TRI_PMV, uplo, trans, diag, n, ap, NULL, x, incx);
}

void dtpsv_(const char *uplo, const char *trans, const char *diag, const int *n, const double *ap,
            double *x, const int *incx)
{
    tri_common(TRI_PSV, uplo, trans, diag, n, ap, NULL, x, incx);
}

/* ---- DSPMV: y := alpha*A*x + beta*y, packed symmetric A (netlib DSPMV, the
 * definition blas_level2/pmv.cc forwards to).  INFO: 1 uplo, 2 n, 6 incx,
 * 9 incy ---- */
void dspmv_(const char *uplo, const int *n, const double *alpha, const double *ap, const double *x,
            const int *incx, const double *beta, double *y, const int *incy)
{
    gb_enter();
    GB_COUNT(calls, 1);
    int info = 0;
    if (!gb_lsame(*uplo, 'U') && !gb_lsame(*uplo, 'L')) info = 1;
    else if (*n < 0) info = 2;
    else if (*incx == 0) info = 6;
    else if (*incy == 0) info = 9;
    if (info) {
        gb_xerbla("DSPMV ", info);
        gb_leave();
        return;
    }
    if (*n == 0 || (*alpha == 0.0 && *beta == 1.0)) {
        gb_leave();
        return;
    }
    const double nn = *n, tri = nn * (nn + 1) / 2;
    if (gb_host_route(GB_RK_DSYMV, gb_opt.cutoff > 0 && *n < gb_opt.cutoff, 2 * nn * nn,
                      8.0 * (tri + 3 * nn), ap, x, y)) {
        void (*f)(const char *, const int *, const double *, const double *, const double *,
                  const int *, const double *, double *, const int *) = gb_host_fn("dspmv_");
        GB_COUNT(fallthrough, 1);
        f(uplo, n, alpha, ap, x, incx, beta, y, incy);
        gb_leave();
        return;
    }
    gb_require_device();
    void *s = gpublas_dev_stream();
    gb_operand ops[3];
    memset(ops, 0, sizeof ops);
    if (*alpha != 0.0) {
        gb_acquire_packed(&ops[0], ap, *n, 8, GB_IN, s);
        gb_acquire_vector(&ops[1], x, *n, *incx, 8, GB_IN, s);
    }
    gb_acquire_vector(&ops[2], y, *n, *incy, 8, *beta == 0.0 ? GB_OUT : GB_INOUT, s);
    gb_check(gpublas_dev_dspmv(s, *uplo, *n, *alpha, (const double *)ops[0].dev,
                               (const double *)ops[1].dev, *incx, *beta, (double *)ops[2].dev,
                               *incy),
             "dspmv launch");
    gb_finish(s, ops, 3);
    GB_COUNT(device_calls, 1);
    gb_leave();
}

/* ---- DGBMV: y := alpha*op(A)*x + beta*y, general band A (netlib DGBMV, the
 * definition blas_level2/gbmv.cc forwards to).  INFO: 1 trans, 2 m, 3 n,
 * 4 kl, 5 ku, 8 lda, 10 incx, 13 incy ---- */
void dgbmv_(const char *trans, const int *m, const int *n, const int *kl, const int *ku,
            const double *alpha, const double *a, const int *lda, const double *x, const int *incx,
            const double *beta, double *y, const int *incy)
{
    gb_enter();
    GB_COUNT(calls, 1);
    int info = 0;
    if (!gb_lsame(*trans, 'N') && !gb_lsame(*trans, 'T') && !gb_lsame(*trans, 'C')) info = 1;
    else if (*m < 0) info = 2;
    else if (*n < 0) info = 3;
    else if (*kl < 0) info = 4;
    else if (*ku < 0) info = 5;
    else if (*lda < *kl + *ku + 1) info = 8;
    else if (*incx == 0) info = 10;
    else if (*incy == 0) info = 13;
    if (info) {
        gb_xerbla("DGBMV ", info);
        gb_leave();
        return;
    }
    if (*m == 0 || *n == 0 || (*alpha == 0.0 && *beta == 1.0)) {
        gb_leave();
        return;
    }
    const int nt = gb_lsame(*trans, 'N');
    const int lenx = nt ? *n : *m, leny = nt ? *m : *n, w = *kl + *ku + 1;
    const double band = (double)w * *n;
    if (gb_host_route(GB_RK_DGEMV,
                      gb_opt.cutoff > 0 && (*m < gb_opt.cutoff || *n < gb_opt.cutoff), 2 * band,
                      8.0 * (band + lenx + leny), a, x, y)) {
        void (*f)(const char *, const int *, const int *, const int *, const int *, const double *,
                  const double *, const int *, const double *, const int *, const double *,
                  double *, const int *) = gb_host_fn("dgbmv_");
        GB_COUNT(fallthrough, 1);
        f(trans, m, n, kl, ku, alpha, a, lda, x, incx, beta, y, incy);
        gb_leave();
        return;
    }
    gb_require_device();
    void *s = gpublas_dev_stream();
    gb_operand ops[3];
    memset(ops, 0, sizeof ops);
    if (*alpha != 0.0) {
        gb_acquire_matrix(&ops[0], a, *lda, w, *n, 8, GB_IN, s);
        gb_acquire_vector(&ops[1], x, lenx, *incx, 8, GB_IN, s);
    }
    gb_acquire_vector(&ops[2], y, leny, *incy, 8, *beta == 0.0 ? GB_OUT : GB_INOUT, s);
    gb_check(gpublas_dev_dgbmv(s, *trans, *m, *n, *kl, *ku, *alpha, (const double *)ops[0].dev,
                               *lda, (const double *)ops[1].dev, *incx, *beta,
                               (double *)ops[2].dev, *incy),
             "dgbmv launch");
    gb_finish(s, ops, 3);
    GB_COUNT(device_calls, 1);
    gb_leave();
}

/* ---- DSBMV: y := alpha*A*x + beta*y, symmetric band A (netlib DSBMV, the
 * definition blas_level2/sbmv.cc forwards to).  INFO: 1 uplo, 2 n, 3 k,
 * 6 lda, 8 incx, 11 incy ---- */
void dsbmv_(const char *uplo, const int *n, const int *k, const double *alpha, const double *a,
            const int *lda, const double *x, const int *incx, const double *beta, double *y,
            const int *incy)
{
    gb_enter();
    GB_COUNT(calls, 1);
    int info = 0;
    if (!gb_lsame(*uplo, 'U') && !gb_lsame(*uplo, 'L')) info = 1;
    else if (*n < 0) info = 2;
    else if (*k < 0) info = 3;
    else if (*lda < *k + 1) info = 6;
    else if (*incx == 0) info = 8;
    else if (*incy == 0) info = 11;
    if (info) {
        gb_xerbla("DSBMV ", info);
        gb_leave();
        return;
    }
    if (*n == 0 || (*alpha == 0.0 && *beta == 1.0)) {
        gb_leave();
        return;
    }
    const double band = (double)(*k + 1) * *n;
    if (gb_host_route(GB_RK_DSYMV, gb_opt.cutoff > 0 && *n < gb_opt.cutoff, 4 * band,
                      8.0 * (band + 3.0 * *n), a, x, y)) {
        void (*f)(const char *, const int *, const int *, const double *, const double *,
                  const int *, const double *, const int *, const double *, double *,
                  const int *) = gb_host_fn("dsbmv_");
        GB_COUNT(fallthrough, 1);
        f(uplo, n, k, alpha, a, lda, x, incx, beta, y, incy);
        gb_leave();
        return;
    }
    gb_require_device();
    void *s = gpublas_dev_stream();
    gb_operand ops[3];
    memset(ops, 0, sizeof ops);
    if (*alpha != 0.0) {
        gb_acquire_matrix(&ops[0], a, *lda, *k + 1, *n, 8, GB_IN, s);
        gb_acquire_vector(&ops[1], x, *n, *incx, 8, GB_IN, s);
    }
    gb_acquire_vector(&ops[2], y, *n, *incy, 8, *beta == 0.0 ? GB_OUT : GB_INOUT, s);
    gb_check(gpublas_dev_dsbmv(s, *uplo, *n, *k, *alpha, (const double *)ops[0].dev, *lda,
                               (const double *)ops[1].dev, *incx, *beta, (double *)ops[2].dev,
                               *incy),
             "dsbmv launch");
    gb_finish(s, ops, 3);
    GB_COUNT(device_calls, 1);
    gb_leave();
}

/* ---- DTBMV / DTBSV: x := op(A)*x and the solve, triangular band A (netlib
 * DTBMV, DTBSV; the definitions blas_level2/tbmv.cc, tbsv.cc forward to).
 * INFO: 1 uplo, 2 trans, 3 diag, 4 n, 5 k, 7 lda, 9 incx.  The device solve
 * keeps the unknowns in flight in LDS, at most 8192 of them: a DTBSV with
 * k > 8191 runs on the host BLAS. ---- */
static void tb_common(int solve, const char *uplo, const char *trans, const char *diag,
                      const int *n, const int *k, const double *a, const int *lda, double *x,
                      const int *incx)
{
    gb_enter();
    GB_COUNT(calls, 1);
    const char *name = solve ? "DTBSV " : "DTBMV ";
    int info = 0;
    if (!gb_lsame(*uplo, 'U') && !gb_lsame(*uplo, 'L')) info = 1;
    else if (!gb_lsame(*trans, 'N') && !gb_lsame(*trans, 'T') && !gb_lsame(*trans, 'C')) info = 2;
    else if (!gb_lsame(*diag, 'U') && !gb_lsame(*diag, 'N')) info = 3;
    else if (*n < 0) info = 4;
    else if (*k < 0) info = 5;
    else if (*lda < *k + 1) info = 7;
    else if (*incx == 0) info = 9;
    if (info) {
        gb_xerbla(name, info);
        gb_leave();
        return;
    }
    if (*n == 0) {
        gb_leave();
        return;
    }
    const double band = (double)(*k + 1) * *n;
    const int host = (solve && *k > 8191) ||
                     gb_host_route(GB_RK_DTRSV, gb_opt.cutoff > 0 && *n < gb_opt.cutoff, 2 * band,
                                   8.0 * (band + 2.0 * *n), a, x, NULL);
    if (host) {
        void (*f)(const char *, const char *, const char *, const int *, const int *,
                  const double *, const int *, double *, const int *) =
            gb_host_fn(solve ? "dtbsv_" : "dtbmv_");
        GB_COUNT(fallthrough, 1);
        f(uplo, trans, diag, n, k, a, lda, x, incx);
        gb_leave();
        return;
    }
    gb_require_device();
    void *s = gpublas_dev_stream();
    gb_operand ops[2];
    memset(ops, 0, sizeof ops);
    gb_acquire_matrix(&ops[0], a, *lda, *k + 1, *n, 8, GB_IN, s);
    gb_acquire_vector(&ops[1], x, *n, *incx, 8, GB_INOUT, s);
    const double *da = (const double *)ops[0].dev;
    double *dx = (double *)ops[1].dev;
    gb_check(solve ? gpublas_dev_dtbsv(s, *uplo, *trans, *diag, *n, *k, da, *lda, dx, *incx)
                   : gpublas_dev_dtbmv(s, *uplo, *trans, *diag, *n, *k, da, *lda, dx, *incx),
             "band triangular launch");
    gb_finish(s, ops, 2);
    GB_COUNT(device_calls, 1);
    gb_leave();
}

void dtbmv_(const char *uplo, const char *trans, const char *diag, const int *n, const int *k,
            const double *a, const int *lda, double *x, const int *incx)
{
    tb_common(0, uplo, trans, diag, n, k, a, lda, x, incx);
}

void dtbsv_(const char *uplo, const char *trans, const char *diag, const int *n, const int *k,
            const double *a, const int *lda, double *x, const int *incx)
{
    tb_common(1, uplo, trans, diag, n, k, a, lda, x, incx);
}

/* library-internal names for the crossover measurement (see level3.c) */
#define GB_SELF(f) extern __typeof__(f) gb_self_##f __attribute__((alias(#f), visibility("hidden")));
GB_SELF(dgemv_)
GB_SELF(dger_)
GB_SELF(dtrsv_)
GB_SELF(dsymv_)
GB_SELF(dsyr_)
GB_SELF(daxpy_)
GB_SELF(ddot_)
