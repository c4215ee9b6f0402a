/*
 * serve.c -- the serving loop: DGEMM / SGEMM / DGEMV / left-side DTRSM calls
 * from any number of host threads, on tracked (device-resident) operands,
 * are collected by one engine thread, admitted under an HBM budget and run
 * as grouped persistent launches (one captured hipGraph replay per batch),
 * with several batches in flight at once.
 *
 * No counterpart in the reference, which runs every intercepted call as its
 * own cuBLAS call on one global handle (runtime-blas.c:14, blas_level3/
 * gemm.cc:62-82) with per-call staging (runtime-mem.hpp:84-165).
 *
 * Life of a served call:
 *   1. The calling thread checks that every operand lies inside a tracked
 *      object, waits for conflicting kernels of the per-call path (its own
 *      or other threads' asynchronous calls), pushes a request onto a
 *      lock-free queue and sleeps on a per-thread futex word.
 *   2. The engine drains the queue into a FIFO.  Whenever a batch slot is
 *      free it admits requests in order: a request whose objects have no
 *      HBM mirror yet needs their size under the budget; idle mirrors whose
 *      host copy is current are evicted least-recently-used first; if that
 *      is not enough the request waits until in-flight batches end (or is
 *      let through when nothing is in flight).  Admission uploads what is
 *      stale and pins the objects.  The batch is launched at once -- it does
 *      not wait to fill: requests that arrive while batches run gather in
 *      the FIFO and form the next batch (continuous batching).
 *   3. A batch runs its kinds (DGEMM, SGEMM, DGEMV, DTRSM groups) cheapest
 *      first, each followed by an event; as soon as a kind's event has
 *      fired the engine marks the objects its requests wrote
 *      device-resident (host access faults and downloads), unpins them and
 *      wakes those callers, which return.  The slot is reused when the whole
 *      batch is done.
 * The slot's kernels give every problem the same work decomposition alone
 * and in any mix, so a served result does not depend on its batch.
 */
#include "internal.h"
#include <linux/futex.h>
#include <stdlib.h>
#include <string.h>
#include <sys/syscall.h>
#include <time.h>
#include <unistd.h>

/* request kinds, numbered as the batch slot's kinds */
enum { RQ_DGEMM = 0, RQ_SGEMM = 1, RQ_DGEMV = 2, RQ_DTRSM = 3 };
#define MAX_SLOTS 16
#define SPIN_POLLS 4000
#define LONG_US 150.0   /* calls estimated above this run in batches of their own class */
#define SCAN_WINDOW 256 /* pending requests a batch may pick from */

struct waiter {
    uint32_t done;      /* ticket of the last completed request (futex word) */
    uint32_t sleeping;  /* caller is (about to be) in FUTEX_WAIT */
    uint32_t issued;
    int status;
};

struct req {
    int op;
    char c0, c1, c2, c3;
    int m, n, k, lda, ldb, ldc, incx, incy;
    double alpha, beta;
    const void *host[3];
    struct gb_obj *obj[3];
    int mode[3];
    int nobj;
    int admitted, deferred;
    int is_long;         /* estimated device time above LONG_US */
    uint64_t t_submit, t_launch;  /* monotonic ns */
    struct waiter *w;
    uint32_t ticket;
    struct req *next;
};

static struct {
    pthread_mutex_t mu;
    int running;
    volatile int stop;
    pthread_t th;
    int nslots;
    gpublas_batch *slot[MAX_SLOTS];
    struct req *slot_reqs[MAX_SLOTS];
    int busy[MAX_SLOTS];
    struct req *queue;          /* lock-free LIFO of submissions */
    uint32_t qword;             /* bumped on every submission (futex word) */
    uint32_t sleeping;
    struct req *pend_head, *pend_tail;  /* engine-only FIFO */
} E = {.mu = PTHREAD_MUTEX_INITIALIZER};

static GB_TLS struct waiter *t_waiter = NULL;

static long futex(uint32_t *uaddr, int op, uint32_t val, const struct timespec *ts)
{
    return syscall(SYS_futex, uaddr, op, val, ts, NULL, 0);
}

static void cpu_relax(void) { __builtin_ia32_pause(); }

static uint64_t now_ns(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (uint64_t)t.tv_sec * 1000000000ull + (uint64_t)t.tv_nsec;
}

/* ---- engine ---------------------------------------------------------- */
static int any_busy(void)
{
    for (int i = 0; i < E.nslots; ++i)
        if (E.busy[i]) return 1;
    return 0;
}

static void complete(struct req *r, int status)
{
    struct waiter *w = r->w;
    const uint32_t t = r->ticket;
    /* r lives on the caller's stack: not touched after the release below */
    w->status = status;
    __atomic_store_n(&w->done, t, __ATOMIC_SEQ_CST);
    if (__atomic_load_n(&w->sleeping, __ATOMIC_SEQ_CST)) futex(&w->done, FUTEX_WAKE_PRIVATE, 1, NULL);
}

static void release_objs(struct req *r, int written)
{
    for (int i = 0; i < r->nobj; ++i) {
        if (written && (r->mode[i] & GB_OUT)) gb_obj_device_written(r->obj[i]);
        gb_obj_unref(r->obj[i]);
    }
}

static int admit(struct req *r, void *stream)
{
    if (r->admitted) return 1;
    size_t need = 0;
    for (int i = 0; i < r->nobj; ++i) {
        int dup = 0;
        for (int j = 0; j < i; ++j) dup |= r->obj[j] == r->obj[i];
        if (!dup && !r->obj[i]->dev) need += r->obj[i]->map_size;
    }
    const size_t budget = gb_opt.serve_budget;
    size_t res = __atomic_load_n(&gb_resident_bytes, __ATOMIC_RELAXED);
    if (need && res + need > budget) {
        gb_evict_idle(res + need - budget, r->obj, r->nobj);
        res = __atomic_load_n(&gb_resident_bytes, __ATOMIC_RELAXED);
        if (res + need > budget) {
            if (any_busy()) {
                if (!r->deferred) GB_COUNT(deferred, 1);
                r->deferred = 1;
                return 0;
            }
            GB_COUNT(over_budget, 1);
        }
    }
    for (int i = 0; i < r->nobj; ++i) gb_obj_to_device(r->obj[i], stream, (r->mode[i] & GB_OUT) != 0);
    r->admitted = 1;
    return 1;
}

static char *dptr(const struct req *r, int i)
{
    const struct gb_obj *o = r->obj[i];
    return o->dev + ((const char *)r->host[i] - o->host);
}

static int add_to_batch(gpublas_batch *b, const struct req *r)
{
    switch (r->op) {
    case RQ_DGEMM:
        return gpublas_dev_batch_dgemm(b, r->c0, r->c1, r->m, r->n, r->k, r->alpha,
                                       (const double *)dptr(r, 0), r->lda,
                                       (const double *)dptr(r, 1), r->ldb, r->beta,
                                       (double *)dptr(r, 2), r->ldc);
    case RQ_SGEMM:
        return gpublas_dev_batch_sgemm(b, r->c0, r->c1, r->m, r->n, r->k, (float)r->alpha,
                                       (const float *)dptr(r, 0), r->lda,
                                       (const float *)dptr(r, 1), r->ldb, (float)r->beta,
                                       (float *)dptr(r, 2), r->ldc);
    case RQ_DGEMV:
        return gpublas_dev_batch_dgemv(b, r->c0, r->m, r->n, r->alpha, (const double *)dptr(r, 0),
                                       r->lda, (const double *)dptr(r, 1), r->incx, r->beta,
                                       (double *)dptr(r, 2), r->incy);
    default:
        return gpublas_dev_batch_dtrsm(b, r->c0, r->c1, r->c2, r->c3, r->m, r->n, r->alpha,
                                       (const double *)dptr(r, 0), r->lda, (double *)dptr(r, 1),
                                       r->ldb);
    }
}

/* Complete every request whose kind has finished (a batch's kinds end at
 * different times); free a slot once all of its work is done. */
static int reap(void)
{
    int progress = 0;
    for (int i = 0; i < E.nslots; ++i) {
        if (!E.busy[i]) continue;
        gpublas_batch *b = E.slot[i];
        int kdone[4] = {-1, -1, -1, -1};
        struct req **link = &E.slot_reqs[i];
        while (*link) {
            struct req *r = *link;
            if (kdone[r->op] < 0) kdone[r->op] = gpublas_dev_batch_done_kind(b, r->op);
            if (!kdone[r->op]) {
                link = &r->next;
                continue;
            }
            *link = r->next;
            GB_COUNT(serve_run_ns, now_ns() - r->t_launch);
            release_objs(r, 1);
            complete(r, 0); /* r is gone after this */
            progress = 1;
        }
        if (!E.slot_reqs[i] && gpublas_dev_batch_done(b)) {
            gpublas_dev_batch_clear(b);
            E.busy[i] = 0;
            progress = 1;
        }
    }
    return progress;
}

static int drain_queue(void)
{
    struct req *l = __atomic_exchange_n(&E.queue, NULL, __ATOMIC_ACQ_REL);
    if (!l) return 0;
    struct req *fifo = NULL;
    while (l) { /* LIFO -> FIFO */
        struct req *next = l->next;
        l->next = fifo;
        fifo = l;
        l = next;
    }
    if (E.pend_tail) E.pend_tail->next = fifo;
    else E.pend_head = fifo;
    while (fifo->next) fifo = fifo->next;
    E.pend_tail = fifo;
    return 1;
}

/* Rough device time of a call (microseconds) from its shape: the batch
 * kernels' sustained rates on MI355X.  Only used to keep long calls from
 * holding short ones back (a batch ends when its longest member does). */
static double est_us(const struct req *r)
{
    const double m = r->m, n = r->n, k = r->k;
    switch (r->op) {
    case RQ_DGEMM: return 2.0 * m * n * k / 30e6;
    case RQ_SGEMM: return 2.0 * m * n * k / 60e6;
    case RQ_DGEMV: return 8.0 * m * n / 3e6;
    default: return 0.6 * (m / 16.0) + m * m * 16.0 / 5e6;
    }
}

static int launch_one(void)
{
    int si = -1;
    for (int i = 0; i < E.nslots; ++i)
        if (!E.busy[i]) {
            si = i;
            break;
        }
    if (si < 0 || !E.pend_head) return 0;
    const uint64_t t0 = now_ns();
    gpublas_batch *b = E.slot[si];
    void *st = gpublas_dev_batch_stream(b);
    struct req *list = NULL, **tail = &list;
    int n = 0;
    /* The oldest request always goes; others of the same class (short or
     * long) join it in arrival order, so a long call never delays short
     * ones and short ones never wait behind it. */
    const int cls = E.pend_head->is_long;
    struct req **link = &E.pend_head, *prev = NULL;
    int scanned = 0;
    while (*link && n < gb_opt.serve_batch && scanned++ < SCAN_WINDOW) {
        struct req *r = *link;
        if (r->is_long != cls) {
            prev = r;
            link = &r->next;
            continue;
        }
        if (!admit(r, st)) break;
        const int rc = add_to_batch(b, r);
        if (rc == 1) break; /* slot full: r leads the next batch */
        *link = r->next; /* unlink */
        if (E.pend_tail == r) E.pend_tail = prev;
        r->next = NULL;
        if (rc < 0) { /* not batchable after all: the caller runs it per call */
            release_objs(r, 0);
            complete(r, -1);
            continue;
        }
        *tail = r;
        tail = &r->next;
        ++n;
    }
    if (!n) return 0;
    gb_check(gpublas_dev_batch_launch(b), "serving batch launch");
    E.busy[si] = 1;
    E.slot_reqs[si] = list;
    const uint64_t t1 = now_ns();
    for (struct req *r = list; r; r = r->next) {
        r->t_launch = t1;
        GB_COUNT(serve_wait_ns, t1 - r->t_submit);
    }
    GB_COUNT(serve_launch_ns, t1 - t0);
    GB_COUNT(batches, 1);
    GB_COUNT(served, n);
    uint64_t mx = __atomic_load_n(&gb_stats.max_batch, __ATOMIC_RELAXED);
    while ((uint64_t)n > mx && !__atomic_compare_exchange_n(&gb_stats.max_batch, &mx, (uint64_t)n,
                                                           0, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
    }
    return 1;
}

static void *engine_main(void *arg)
{
    (void)arg;
    gb_guard = 1; /* allocations made on this thread are never tracked */
    unsigned idle = 0;
    for (;;) {
        int progress = reap();
        progress |= drain_queue();
        while (launch_one()) progress = 1;
        if (progress) {
            idle = 0;
            continue;
        }
        const int busy = any_busy();
        if (!busy && !E.pend_head && E.stop) break;
        if (busy || E.pend_head || ++idle < 2000) {
            cpu_relax();
            continue;
        }
        /* nothing in flight or waiting: sleep until a submission */
        const uint32_t seen = __atomic_load_n(&E.qword, __ATOMIC_SEQ_CST);
        __atomic_store_n(&E.sleeping, 1, __ATOMIC_SEQ_CST);
        if (!__atomic_load_n(&E.queue, __ATOMIC_SEQ_CST) && !E.stop) {
            struct timespec ts = {0, 100 * 1000 * 1000};
            futex(&E.qword, FUTEX_WAIT_PRIVATE, seen, &ts);
        }
        __atomic_store_n(&E.sleeping, 0, __ATOMIC_SEQ_CST);
        idle = 0;
    }
    return NULL;
}

static int engine_start(void)
{
    if (__atomic_load_n(&E.running, __ATOMIC_ACQUIRE)) return 1;
    pthread_mutex_lock(&E.mu);
    if (!E.running) {
        gb_require_device();
        if (gb_opt.tune_file[0] && gpublas_dev_tune_gemm(gb_opt.tune_file, 0) != 0)
            gb_log(2, "libgpublas: tile-shape table '%s' unusable; using the defaults\n",
                   gb_opt.tune_file);
        E.nslots = gb_opt.serve_slots;
        int ok = 1;
        for (int i = 0; i < E.nslots && ok; ++i) {
            E.slot[i] = gpublas_dev_batch_create(gb_opt.serve_graph);
            ok = E.slot[i] != NULL;
            E.busy[i] = 0;
        }
        E.stop = 0;
        E.pend_head = E.pend_tail = NULL;
        if (ok && pthread_create(&E.th, NULL, engine_main, NULL) == 0) {
            __atomic_store_n(&E.running, 1, __ATOMIC_RELEASE);
        } else {
            for (int i = 0; i < E.nslots; ++i)
                if (E.slot[i]) gpublas_dev_batch_destroy(E.slot[i]);
            gb_log(2, "libgpublas: serving loop could not start; calls run one by one\n");
        }
    }
    const int r = E.running;
    pthread_mutex_unlock(&E.mu);
    return r;
}

void gb_serve_stop(void)
{
    pthread_mutex_lock(&E.mu);
    if (E.running) {
        E.stop = 1;
        __atomic_add_fetch(&E.qword, 1, __ATOMIC_SEQ_CST);
        futex(&E.qword, FUTEX_WAKE_PRIVATE, 1, NULL);
        pthread_join(E.th, NULL);
        for (int i = 0; i < E.nslots; ++i) {
            gpublas_dev_batch_destroy(E.slot[i]);
            E.slot[i] = NULL;
        }
        __atomic_store_n(&E.running, 0, __ATOMIC_RELEASE);
    }
    pthread_mutex_unlock(&E.mu);
}

/* ---- callers --------------------------------------------------------- */
static int find_operand(struct req *r, const void *host, size_t ld, size_t rows, size_t cols,
                        size_t es, int mode)
{
    const size_t span = ((cols - 1) * ld + rows) * es;
    struct gb_obj *o = gb_obj_find(host);
    if (!o || (const char *)host + span > o->host + o->size) return 0;
    const int i = r->nobj++;
    r->host[i] = host;
    r->obj[i] = o;
    r->mode[i] = mode;
    return 1;
}

static int vec_operand(struct req *r, const void *host, int n, int inc, int mode)
{
    const size_t ainc = (size_t)(inc < 0 ? -inc : inc);
    const size_t span = 1 + (size_t)(n - 1) * ainc;
    return find_operand(r, host, span, span, 1, 8, mode);
}

/* Submit and wait.  Returns 1 if the engine ran the call. */
static int submit(struct req *r)
{
    if (!engine_start()) return 0;
    if (!t_waiter) {
        t_waiter = calloc(1, sizeof *t_waiter); /* kept for the thread's life */
        if (!t_waiter) return 0;
    }
    struct waiter *w = t_waiter;
    r->w = w;
    r->ticket = ++w->issued;
    r->t_submit = now_ns();
    /* kernels of the per-call path (any thread) that this call must follow */
    for (int i = 0; i < r->nobj; ++i) gb_obj_wait_hazards(r->obj[i], (r->mode[i] & GB_OUT) != 0);
    struct req *head = __atomic_load_n(&E.queue, __ATOMIC_RELAXED);
    do {
        r->next = head;
    } while (!__atomic_compare_exchange_n(&E.queue, &head, r, 1, __ATOMIC_RELEASE,
                                          __ATOMIC_RELAXED));
    __atomic_add_fetch(&E.qword, 1, __ATOMIC_SEQ_CST);
    if (__atomic_load_n(&E.sleeping, __ATOMIC_SEQ_CST)) futex(&E.qword, FUTEX_WAKE_PRIVATE, 1, NULL);
    const uint32_t t = r->ticket;
    for (int spin = 0; spin < SPIN_POLLS; ++spin) {
        if (__atomic_load_n(&w->done, __ATOMIC_ACQUIRE) == t) goto done;
        cpu_relax();
    }
    __atomic_store_n(&w->sleeping, 1, __ATOMIC_SEQ_CST);
    for (;;) {
        const uint32_t v = __atomic_load_n(&w->done, __ATOMIC_SEQ_CST);
        if (v == t) break;
        futex(&w->done, FUTEX_WAIT_PRIVATE, v, NULL);
    }
    __atomic_store_n(&w->sleeping, 0, __ATOMIC_SEQ_CST);
done:
    return w->status == 0;
}

int gb_serve_dgemm(char ta, char tb, int m, int n, int k, double alpha, const double *a, int lda,
                   const double *b, int ldb, double beta, double *c, int ldc)
{
    if (gb_opt.split > 0 || alpha == 0 || k == 0) return 0;
    struct req r;
    memset(&r, 0, sizeof r);
    const int nota = gb_lsame(ta, 'N'), notb = gb_lsame(tb, 'N');
    if (!find_operand(&r, a, lda, nota ? m : k, nota ? k : m, 8, GB_IN) ||
        !find_operand(&r, b, ldb, notb ? k : n, notb ? n : k, 8, GB_IN) ||
        !find_operand(&r, c, ldc, m, n, 8, beta == 0 ? GB_OUT : GB_INOUT))
        return 0;
    r.op = RQ_DGEMM;
    r.c0 = nota ? 'N' : 'T';
    r.c1 = notb ? 'N' : 'T';
    r.m = m, r.n = n, r.k = k, r.lda = lda, r.ldb = ldb, r.ldc = ldc;
    r.alpha = alpha, r.beta = beta;
    r.is_long = est_us(&r) > LONG_US;
    return submit(&r);
}

int gb_serve_sgemm(char ta, char tb, int m, int n, int k, float alpha, const float *a, int lda,
                   const float *b, int ldb, float beta, float *c, int ldc)
{
    if (alpha == 0 || k == 0) return 0;
    struct req r;
    memset(&r, 0, sizeof r);
    const int nota = gb_lsame(ta, 'N'), notb = gb_lsame(tb, 'N');
    if (!find_operand(&r, a, lda, nota ? m : k, nota ? k : m, 4, GB_IN) ||
        !find_operand(&r, b, ldb, notb ? k : n, notb ? n : k, 4, GB_IN) ||
        !find_operand(&r, c, ldc, m, n, 4, beta == 0 ? GB_OUT : GB_INOUT))
        return 0;
    r.op = RQ_SGEMM;
    r.c0 = nota ? 'N' : 'T';
    r.c1 = notb ? 'N' : 'T';
    r.m = m, r.n = n, r.k = k, r.lda = lda, r.ldb = ldb, r.ldc = ldc;
    r.alpha = alpha, r.beta = beta;
    r.is_long = est_us(&r) > LONG_US;
    return submit(&r);
}

int gb_serve_dgemv(char trans, int m, int n, double alpha, const double *a, int lda,
                   const double *x, int incx, double beta, double *y, int incy)
{
    if (alpha == 0) return 0;
    struct req r;
    memset(&r, 0, sizeof r);
    const int notr = gb_lsame(trans, 'N');
    const int lenx = notr ? n : m, leny = notr ? m : n;
    /* a strided output has gaps: it is read as well as written */
    const int ymode = (beta == 0 && (incy == 1 || incy == -1)) ? GB_OUT : GB_INOUT;
    if (!find_operand(&r, a, lda, m, n, 8, GB_IN) || !vec_operand(&r, x, lenx, incx, GB_IN) ||
        !vec_operand(&r, y, leny, incy, ymode))
        return 0;
    r.op = RQ_DGEMV;
    r.c0 = notr ? 'N' : 'T';
    r.m = m, r.n = n, r.lda = lda, r.incx = incx, r.incy = incy;
    r.alpha = alpha, r.beta = beta;
    r.is_long = est_us(&r) > LONG_US;
    return submit(&r);
}

int gb_serve_dtrsm(char side, char uplo, char transa, char diag, int m, int n, double alpha,
                   const double *a, int lda, double *b, int ldb)
{
    if (!gb_lsame(side, 'L') || alpha == 0 || m > 1024) return 0; /* batch kernel: m <= 1024 */
    struct req r;
    memset(&r, 0, sizeof r);
    if (!find_operand(&r, a, lda, m, m, 8, GB_IN) || !find_operand(&r, b, ldb, m, n, 8, GB_INOUT))
        return 0;
    r.op = RQ_DTRSM;
    r.c0 = 'L';
    r.c1 = gb_lsame(uplo, 'U') ? 'U' : 'L';
    r.c2 = gb_lsame(transa, 'N') ? 'N' : 'T';
    r.c3 = gb_lsame(diag, 'U') ? 'U' : 'N';
    r.m = m, r.n = n, r.lda = lda, r.ldb = ldb;
    r.alpha = alpha;
    r.is_long = est_us(&r) > LONG_US;
    return submit(&r);
}
