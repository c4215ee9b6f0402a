/*
 * serve_bench.c -- serving trace: many host threads issue intercepted BLAS
 * calls (Fortran ABI of libgpublas.so) with Poisson arrivals; per-call
 * latency is measured from the scheduled arrival to the result being
 * complete on the device, so queueing in front of a busy caller counts.
 *
 * Mix (each thread draws its own calls from a seeded SplitMix64 stream):
 *   60% GEMM, half DGEMM / half SGEMM, M, N, K each log-uniform in
 *       [32, 2048], transa / transb random, alpha 1, beta 0;
 *   30% DGEMV, m, n log-uniform in [256, 16384], trans random, alpha 1,
 *       beta 0, unit strides;
 *   10% DTRSM, side L, lower, no-trans, non-unit, m log-uniform in
 *       [64, 1024], nrhs 256, A diagonally dominant (diag 1024 + U[0,1],
 *       off-diagonal U[-1,1]), alpha 1024 (keeps the right-hand sides of
 *       repeated solves at unit magnitude).
 * All entries uniform[-1,1].  Operands are tracked objects made resident
 * before timing: shared read-only pools (A/B of GEMM, A/x of GEMV, A of
 * TRSM; each call reads a random window) and per-thread outputs.
 *
 * Modes:
 *   plain  -- serving loop off: every call is its own launch (the library's
 *             per-call path) and the caller waits for its kernel
 *             (gpublas_sync) before taking the next arrival;
 *   serve  -- serving loop on (GPUBLAS_OPTIONS serve=..., serve_graph, ...
 *             are passed through --options).
 * --verify k checks every k-th timed call of each thread against a host BLAS
 * (--verify-blas) after its latency is taken; the per-kind counts, failures
 * and worst normwise errors are added to the output, and the exit status is
 * 4 if any check failed.  Throughput measured with checks on includes them.
 * Prints one JSON object.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "gpublas.h"

typedef void (*gemm_t)(const char *, const char *, const int *, const int *, const int *,
                       const void *, const void *, const int *, const void *, const int *,
                       const void *, void *, const int *);
typedef void (*gemv_t)(const char *, const int *, const int *, const double *, const double *,
                       const int *, const double *, const int *, const double *, double *,
                       const int *);
typedef void (*trsm_t)(const char *, const char *, const char *, const char *, const int *,
                       const int *, const double *, const double *, const int *, double *,
                       const int *);

typedef gpublas_stats_t stats_t; /* the library fills the whole struct */

static gemm_t p_dgemm, p_sgemm;
static gemv_t p_dgemv;
static trsm_t p_dtrsm;
static void *(*p_malloc)(size_t);
static void *(*p_devptr)(const void *);
static int (*p_setopt)(const char *);
static void (*p_stats)(stats_t *);
static void (*p_reset)(void);
static void (*p_sync)(void);
static int (*p_fill)(void *, uint64_t, int64_t, double *);
static int (*p_fill32)(void *, uint64_t, int64_t, float *);
static int (*p_devsync)(void *);

enum { OP_DGEMM, OP_SGEMM, OP_DGEMV, OP_DTRSM, NOPS };
static const char *op_name[NOPS] = {"dgemm", "sgemm", "dgemv", "dtrsm"};

#define GP 4096     /* GEMM pool: GP x GP, windows up to 2048 x 2048 */
#define VP 16384    /* GEMV pool: VP x VP */
#define TP 1024     /* TRSM matrix */
#define NRHS 256

static double *g_poold, *g_poolx, *g_poolv, *g_poolt;
static float *g_pools;

static int g_threads = 16, g_calls = 20000, g_warmup = 16, g_plain = 0;
static double g_rate = 0.0; /* calls/s over all threads; 0 = closed loop */
static uint64_t g_seed = 777;

typedef struct {
    int id;
    double *cd, *y, *bt;
    float *cs;
    double *lat;   /* per timed call, seconds */
    int *op;
    double t_end;
} worker_t;

static double now(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

static uint64_t mix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double unif(uint64_t *s) { return (double)(mix64(s) >> 11) * 0x1.0p-53; }
static int logu(uint64_t *s, int lo, int hi)
{
    const double v = exp(log((double)lo) + unif(s) * (log((double)hi) - log((double)lo)));
    int r = (int)(v + 0.5);
    return r < lo ? lo : r > hi ? hi : r;
}

static void *sym(void *h, const char *name)
{
    void *f = dlsym(h, name);
    if (!f) {
        fprintf(stderr, "serve_bench: no symbol %s\n", name);
        exit(2);
    }
    return f;
}

/* tracked, device-resident, filled on the device */
static void *resident(size_t elems, int single, uint64_t seed)
{
    void *h = p_malloc(elems * (single ? 4 : 8));
    if (!h) {
        fprintf(stderr, "serve_bench: gpublas_malloc failed\n");
        exit(3);
    }
    void *d = p_devptr(h);
    if (single) p_fill32(NULL, seed, (int64_t)elems, (float *)d);
    else p_fill(NULL, seed, (int64_t)elems, (double *)d);
    p_devsync(NULL);
    return h;
}

/* One drawn call: everything needed to run it and to check it afterwards. */
typedef struct {
    int op, m, n, k, lenx;
    char ta, tb;
    size_t ia, ib, ix;
} call_t;

static void draw_call(uint64_t *s, call_t *c)
{
    const double u = unif(s);
    memset(c, 0, sizeof *c);
    if (u < 0.6) {
        const int single = unif(s) < 0.5;
        c->m = logu(s, 32, 2048), c->n = logu(s, 32, 2048), c->k = logu(s, 32, 2048);
        c->ta = unif(s) < 0.5 ? 'N' : 'T', c->tb = unif(s) < 0.5 ? 'N' : 'T';
        const int ra = c->ta == 'N' ? c->m : c->k, ca = c->ta == 'N' ? c->k : c->m;
        const int rb = c->tb == 'N' ? c->k : c->n, cb = c->tb == 'N' ? c->n : c->k;
        c->ia = (size_t)(unif(s) * (GP - ra)) + (size_t)(unif(s) * (GP - ca)) * GP;
        c->ib = (size_t)(unif(s) * (GP - rb)) + (size_t)(unif(s) * (GP - cb)) * GP;
        c->op = single ? OP_SGEMM : OP_DGEMM;
    } else if (u < 0.9) {
        c->m = logu(s, 256, VP), c->n = logu(s, 256, VP);
        c->ta = unif(s) < 0.5 ? 'N' : 'T';
        c->ia = (size_t)(unif(s) * (VP - c->m)) + (size_t)(unif(s) * (VP - c->n)) * VP;
        c->lenx = c->ta == 'N' ? c->n : c->m;
        c->ix = (size_t)(unif(s) * (VP - c->lenx));
        c->op = OP_DGEMV;
    } else {
        c->m = logu(s, 64, TP), c->n = NRHS;
        c->op = OP_DTRSM;
    }
}

static const double TRSM_ALPHA = 1024.0;

static void run_call(worker_t *w, const call_t *c)
{
    const double one = 1.0, zero = 0.0;
    const float onef = 1.0f, zerof = 0.0f;
    const int ld = c->op == OP_DGEMV ? VP : c->op == OP_DTRSM ? TP : GP, inc = 1, ldc = c->m;
    switch (c->op) {
    case OP_SGEMM:
        p_sgemm(&c->ta, &c->tb, &c->m, &c->n, &c->k, &onef, g_pools + c->ia, &ld, g_pools + c->ib, &ld,
                &zerof, w->cs, &ldc);
        break;
    case OP_DGEMM:
        p_dgemm(&c->ta, &c->tb, &c->m, &c->n, &c->k, &one, g_poold + c->ia, &ld, g_poold + c->ib, &ld,
                &zero, w->cd, &ldc);
        break;
    case OP_DGEMV:
        p_dgemv(&c->ta, &c->m, &c->n, &one, g_poolv + c->ia, &ld, g_poolx + c->ix, &inc, &zero, w->y,
                &inc);
        break;
    default:
        p_dtrsm("L", "L", "N", "N", &c->m, &c->n, &TRSM_ALPHA, g_poolt, &ld, w->bt, &ld);
        break;
    }
    if (g_plain) p_sync();
}

/* ---- verification against a host BLAS (--verify k --verify-blas lib) ----
 * Every k-th timed call of each thread is checked once it has returned,
 * before the thread's next call reuses its output buffer:
 *   GEMM  ||C - Cref||_F / (||op(A)||_F ||op(B)||_F k eps) <= 4, Cref from
 *         the host DGEMM in FP64 on the same (FP32-widened) operands;
 *   GEMV  ||y - yref|| / (||A||_F ||x|| len eps) <= 4, yref from host DGEMV;
 *   TRSM  ||A X - alpha B0||_F / (||A||_F ||X||_F m eps) <= 4, A X from host
 *         DTRMM; B0 is copied before the call.
 * eps is that of the call's precision.  Reading the tracked outputs and pool
 * windows from the host goes through the library's coherence (a download). */
typedef void (*hgemm_t)(const char *, const char *, const int *, const int *, const int *,
                        const double *, const double *, const int *, const double *, const int *,
                        const double *, double *, const int *);
typedef void (*htrmm_t)(const char *, const char *, const char *, const char *, const int *,
                        const int *, const double *, const double *, const int *, double *,
                        const int *);
static int g_verify = 0;
static hgemm_t h_dgemm;
static gemv_t h_dgemv;
static htrmm_t h_dtrmm;
static const double VERIFY_BAR = 4.0;

typedef struct {
    double *a, *b, *r; /* scratch: widened operands / saved B0, reference result */
    int checked[NOPS], failed[NOPS];
    double worst[NOPS];
} verify_t;
static verify_t *g_ver;

static double fro(const double *a, int rows, int cols, int ld)
{
    double s = 0;
    for (int j = 0; j < cols; ++j)
        for (int i = 0; i < rows; ++i) s += a[i + (size_t)j * ld] * a[i + (size_t)j * ld];
    return sqrt(s);
}

static void widen(const float *src, int rows, int cols, int ld, double *dst)
{
    for (int j = 0; j < cols; ++j)
        for (int i = 0; i < rows; ++i) dst[i + (size_t)j * rows] = src[i + (size_t)j * ld];
}

static void verify_call(worker_t *w, const call_t *c, const double *b0, verify_t *v)
{
    const double one = 1.0, zero = 0.0;
    double err = 0;
    if (c->op == OP_DGEMM || c->op == OP_SGEMM) {
        const int ra = c->ta == 'N' ? c->m : c->k, ca = c->ta == 'N' ? c->k : c->m;
        const int rb = c->tb == 'N' ? c->k : c->n, cb = c->tb == 'N' ? c->n : c->k;
        const double *A, *B;
        int lda, ldb;
        double diff = 0;
        if (c->op == OP_DGEMM) {
            A = g_poold + c->ia, B = g_poold + c->ib, lda = ldb = GP;
        } else {
            widen(g_pools + c->ia, ra, ca, GP, v->a);
            widen(g_pools + c->ib, rb, cb, GP, v->b);
            A = v->a, B = v->b, lda = ra, ldb = rb;
        }
        h_dgemm(&c->ta, &c->tb, &c->m, &c->n, &c->k, &one, A, &lda, B, &ldb, &zero, v->r, &c->m);
        for (size_t i = 0; i < (size_t)c->m * c->n; ++i) {
            const double g = c->op == OP_DGEMM ? w->cd[i] : (double)w->cs[i];
            diff += (g - v->r[i]) * (g - v->r[i]);
        }
        const double eps = c->op == OP_DGEMM ? 0x1.0p-52 : 0x1.0p-23;
        const double den = fro(A, ra, ca, lda) * fro(B, rb, cb, ldb) * c->k * eps;
        err = den > 0 ? sqrt(diff) / den : sqrt(diff);
    } else if (c->op == OP_DGEMV) {
        const int inc = 1, ld = VP, leny = c->ta == 'N' ? c->m : c->n;
        h_dgemv(&c->ta, &c->m, &c->n, &one, g_poolv + c->ia, &ld, g_poolx + c->ix, &inc, &zero,
                v->r, &inc);
        double diff = 0;
        for (int i = 0; i < leny; ++i) diff += (w->y[i] - v->r[i]) * (w->y[i] - v->r[i]);
        const double den = fro(g_poolv + c->ia, c->m, c->n, VP) * fro(g_poolx + c->ix, c->lenx, 1, 1) *
                           c->lenx * 0x1.0p-52;
        err = den > 0 ? sqrt(diff) / den : sqrt(diff);
    } else {
        const int m = c->m, n = c->n, ld = TP;
        for (int j = 0; j < n; ++j)
            memcpy(v->r + (size_t)j * m, w->bt + (size_t)j * TP, sizeof(double) * (size_t)m);
        h_dtrmm("L", "L", "N", "N", &m, &n, &one, g_poolt, &ld, v->r, &m);
        double diff = 0, ta = 0;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i) {
                const double d = v->r[i + (size_t)j * m] - TRSM_ALPHA * b0[i + (size_t)j * m];
                diff += d * d;
            }
        for (int j = 0; j < m; ++j)
            for (int i = j; i < m; ++i) ta += g_poolt[i + (size_t)j * TP] * g_poolt[i + (size_t)j * TP];
        const double den = sqrt(ta) * fro(w->bt, m, n, TP) * m * 0x1.0p-52;
        err = den > 0 ? sqrt(diff) / den : sqrt(diff);
    }
    v->checked[c->op]++;
    if (!(err <= VERIFY_BAR)) v->failed[c->op]++;
    if (!(err <= v->worst[c->op])) v->worst[c->op] = err; /* NaN sticks */
}

/* Runs one drawn call; with `check` it also saves what the check needs
 * (the right-hand sides a TRSM overwrites) into the thread's scratch. */
static void one_call(worker_t *w, uint64_t *s, call_t *c, int check)
{
    draw_call(s, c);
    if (check && c->op == OP_DTRSM)
        for (int j = 0; j < c->n; ++j)
            memcpy(g_ver[w->id].b + (size_t)j * c->m, w->bt + (size_t)j * TP,
                   sizeof(double) * (size_t)c->m);
    run_call(w, c);
}

static void *worker_main(void *arg)
{
    worker_t *w = arg;
    uint64_t s = g_seed * 1000003ull + (uint64_t)w->id;
    const int per = g_calls / g_threads;
    call_t c;
    for (int i = 0; i < g_warmup; ++i) one_call(w, &s, &c, 0);
    const double rate_t = g_rate > 0 ? g_rate / g_threads : 0;
    double t_sched = now();
    for (int i = 0; i < per; ++i) {
        if (rate_t > 0) {
            t_sched += -log(1.0 - unif(&s)) / rate_t;
            double t;
            while ((t = now()) < t_sched) {
                const double d = t_sched - t;
                if (d > 200e-6) {
                    struct timespec ts = {0, (long)((d - 100e-6) * 1e9)};
                    nanosleep(&ts, NULL);
                }
            }
        } else {
            t_sched = now();
        }
        const int check = g_verify > 0 && i % g_verify == 0;
        one_call(w, &s, &c, check);
        w->lat[i] = now() - t_sched;
        w->op[i] = c.op;
        if (check) verify_call(w, &c, g_ver[w->id].b, &g_ver[w->id]);
    }
    w->t_end = now();
    return NULL;
}

static int cmpd(const void *a, const void *b)
{
    const double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : x > y;
}

static double pct(const double *v, int n, double p)
{
    if (n <= 0) return 0;
    int i = (int)ceil(p * n) - 1;
    if (i < 0) i = 0;
    if (i >= n) i = n - 1;
    return v[i];
}

int main(int argc, char **argv)
{
    const char *lib = NULL, *opts = "", *vblas = NULL, *vprefix = "";
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "--lib") && i + 1 < argc) lib = argv[++i];
        else if (!strcmp(argv[i], "--threads") && i + 1 < argc) g_threads = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--calls") && i + 1 < argc) g_calls = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--warmup") && i + 1 < argc) g_warmup = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--rate") && i + 1 < argc) g_rate = atof(argv[++i]);
        else if (!strcmp(argv[i], "--seed") && i + 1 < argc) g_seed = strtoull(argv[++i], 0, 10);
        else if (!strcmp(argv[i], "--plain")) g_plain = 1;
        else if (!strcmp(argv[i], "--options") && i + 1 < argc) opts = argv[++i];
        else if (!strcmp(argv[i], "--verify") && i + 1 < argc) g_verify = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--verify-blas") && i + 1 < argc) vblas = argv[++i];
        else if (!strcmp(argv[i], "--verify-prefix") && i + 1 < argc) vprefix = argv[++i];
    }
    if (!lib) {
        fprintf(stderr, "usage: serve_bench --lib libgpublas.so [--plain] [--options o] "
                        "[--rate calls/s] [--threads t] [--calls n] "
                        "[--verify k --verify-blas host_blas.so [--verify-prefix p]]\n");
        return 2;
    }
    void *h = dlopen(lib, RTLD_NOW | RTLD_GLOBAL);
    if (!h) {
        fprintf(stderr, "serve_bench: %s\n", dlerror());
        return 2;
    }
    p_dgemm = (gemm_t)sym(h, "dgemm_");
    p_sgemm = (gemm_t)sym(h, "sgemm_");
    p_dgemv = (gemv_t)sym(h, "dgemv_");
    p_dtrsm = (trsm_t)sym(h, "dtrsm_");
    p_malloc = sym(h, "gpublas_malloc");
    p_devptr = sym(h, "gpublas_device_ptr");
    p_setopt = sym(h, "gpublas_set_options");
    p_stats = sym(h, "gpublas_get_stats");
    p_reset = sym(h, "gpublas_reset_stats");
    p_sync = sym(h, "gpublas_sync");
    p_fill = sym(h, "gpublas_dev_fill_uniform");
    p_fill32 = sym(h, "gpublas_dev_fill_uniform_f32");
    p_devsync = sym(h, "gpublas_dev_sync");
    ((void (*)(int))sym(h, "gpublas_set_cutoff"))(0);
    if (g_verify > 0) {
        if (!vblas) {
            fprintf(stderr, "serve_bench: --verify needs --verify-blas <host BLAS library>\n");
            return 2;
        }
        void *hb = dlopen(vblas, RTLD_NOW | RTLD_LOCAL);
        if (!hb) {
            fprintf(stderr, "serve_bench: %s\n", dlerror());
            return 2;
        }
        char nm[128];
        snprintf(nm, sizeof nm, "%sdgemm_", vprefix);
        h_dgemm = (hgemm_t)sym(hb, nm);
        snprintf(nm, sizeof nm, "%sdgemv_", vprefix);
        h_dgemv = (gemv_t)sym(hb, nm);
        snprintf(nm, sizeof nm, "%sdtrmm_", vprefix);
        h_dtrmm = (htrmm_t)sym(hb, nm);
        /* one host thread per caller: 16 callers each checking on a
         * multi-threaded host BLAS would oversubscribe the cores */
        snprintf(nm, sizeof nm, "%sopenblas_set_num_threads", vprefix);
        void (*setnt)(int) = (void (*)(int))dlsym(hb, nm);
        if (setnt) setnt(1);
    }
    if (g_plain) p_setopt("serve=0");
    else if (p_setopt(*opts ? opts : "serve") != 0) {
        fprintf(stderr, "serve_bench: bad options '%s'\n", opts);
        return 2;
    }
    if (!g_plain && !strstr(opts, "serve")) p_setopt("serve");

    const double t0 = now();
    g_poold = resident((size_t)GP * GP, 0, 11);
    g_pools = resident((size_t)GP * GP, 1, 12);
    g_poolv = resident((size_t)VP * VP, 0, 13);
    g_poolx = resident((size_t)VP, 0, 14);
    g_poolt = resident((size_t)TP * TP, 0, 15);
    for (int i = 0; i < TP; ++i) /* diagonal dominance (host write; uploaded once) */
        g_poolt[i + (size_t)i * TP] = TP + 0.5 * (g_poolt[i + (size_t)i * TP] + 1.0);
    const int per = g_calls / g_threads;
    worker_t *w = calloc((size_t)g_threads, sizeof *w);
    for (int t = 0; t < g_threads; ++t) {
        w[t].id = t;
        w[t].cd = resident((size_t)2048 * 2048, 0, 100 + t);
        w[t].cs = resident((size_t)2048 * 2048, 1, 200 + t);
        w[t].y = resident((size_t)VP, 0, 300 + t);
        w[t].bt = resident((size_t)TP * NRHS, 0, 400 + t);
        w[t].lat = calloc((size_t)per, sizeof(double));
        if (g_verify > 0) {
            if (!g_ver) g_ver = calloc((size_t)g_threads, sizeof *g_ver);
            const size_t big = (size_t)2048 * 2048;
            g_ver[t].a = malloc(big * sizeof(double));
            g_ver[t].b = malloc(big * sizeof(double));
            g_ver[t].r = malloc((big > VP ? big : VP) * sizeof(double));
            if (!g_ver[t].a || !g_ver[t].b || !g_ver[t].r) {
                fprintf(stderr, "serve_bench: out of host memory for --verify\n");
                return 3;
            }
        }
        w[t].op = calloc((size_t)per, sizeof(int));
    }
    const double t_setup = now() - t0;
    p_reset();
    pthread_t *th = calloc((size_t)g_threads, sizeof *th);
    const double t1 = now();
    for (int t = 0; t < g_threads; ++t) pthread_create(&th[t], NULL, worker_main, &w[t]);
    for (int t = 0; t < g_threads; ++t) pthread_join(th[t], NULL);
    double t_last = t1;
    for (int t = 0; t < g_threads; ++t)
        if (w[t].t_end > t_last) t_last = w[t].t_end;
    const double wall = t_last - t1;
    stats_t st;
    p_stats(&st);

    const int total = per * g_threads;
    double *all = malloc(sizeof(double) * (size_t)total);
    double *byop[NOPS];
    int nop[NOPS] = {0};
    for (int o = 0; o < NOPS; ++o) byop[o] = malloc(sizeof(double) * (size_t)total);
    int q = 0;
    for (int t = 0; t < g_threads; ++t)
        for (int i = 0; i < per; ++i) {
            all[q++] = w[t].lat[i];
            const int o = w[t].op[i];
            byop[o][nop[o]++] = w[t].lat[i];
        }
    qsort(all, (size_t)total, sizeof(double), cmpd);
    printf("{\"mode\": \"%s\", \"options\": \"%s\", \"threads\": %d, \"calls\": %d, "
           "\"offered_rate\": %.1f, \"wall_s\": %.4f, \"calls_per_s\": %.1f, "
           "\"p50_us\": %.1f, \"p90_us\": %.1f, \"p99_us\": %.1f, \"max_us\": %.1f",
           g_plain ? "plain" : "serve", g_plain ? "" : opts, g_threads, total, g_rate, wall,
           total / wall, 1e6 * pct(all, total, 0.5), 1e6 * pct(all, total, 0.9),
           1e6 * pct(all, total, 0.99), 1e6 * all[total - 1]);
    for (int o = 0; o < NOPS; ++o) {
        qsort(byop[o], (size_t)nop[o], sizeof(double), cmpd);
        printf(", \"%s\": {\"n\": %d, \"p50_us\": %.1f, \"p99_us\": %.1f}", op_name[o], nop[o],
               1e6 * pct(byop[o], nop[o], 0.5), 1e6 * pct(byop[o], nop[o], 0.99));
    }
    const double ns = st.served ? (double)st.served : 1.0;
    printf(", \"engine_wait_us\": %.1f, \"engine_run_us\": %.1f, \"engine_launch_us_per_batch\": %.1f",
           st.serve_wait_ns / ns / 1e3, st.serve_run_ns / ns / 1e3,
           st.batches ? st.serve_launch_ns / (double)st.batches / 1e3 : 0.0);
    printf(", \"served\": %llu, \"batches\": %llu, \"mean_batch\": %.2f, \"max_batch\": %llu, "
           "\"deferred\": %llu, \"evictions\": %llu, \"resident_bytes\": %llu, \"setup_s\": %.2f",
           (unsigned long long)st.served, (unsigned long long)st.batches,
           st.batches ? (double)st.served / (double)st.batches : 0.0,
           (unsigned long long)st.max_batch, (unsigned long long)st.deferred,
           (unsigned long long)st.evictions, (unsigned long long)st.resident_bytes, t_setup);
    int bad = 0;
    if (g_verify > 0) {
        /* the closing brace of the object moves behind the verify summary */
        printf(", \"verify_every\": %d, \"verify_bar\": %.1f", g_verify, VERIFY_BAR);
        for (int o = 0; o < NOPS; ++o) {
            int ck = 0, fl = 0;
            double worst = 0;
            for (int t = 0; t < g_threads; ++t) {
                ck += g_ver[t].checked[o];
                fl += g_ver[t].failed[o];
                if (!(g_ver[t].worst[o] <= worst)) worst = g_ver[t].worst[o];
            }
            bad += fl;
            printf(", \"verify_%s\": {\"checked\": %d, \"failed\": %d, \"max_err\": %.4g}", op_name[o],
                   ck, fl, worst);
        }
    }
    printf("}\n");
    fflush(stdout);
    return bad ? 4 : 0;
}
