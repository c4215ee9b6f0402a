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

typedef void (*gemm_t)(const char *, const char *, const int *, const int *, const int *,
                       const void *, const void *, const int *, const void *, const int *,
                       const void *, void *, const int *);
typedef void (*gemv_t)(const char *, const int *, const int *, const double *, const double *,
                       const int *, const double *, const int *, const double *, double *,
                       const int *);
typedef void (*trsm_t)(const char *, const char *, const char *, const char *, const int *,
                       const int *, const double *, const double *, const int *, double *,
                       const int *);

typedef struct {
    uint64_t calls, device_calls, fallthrough, hits, misses, h2d_bytes, d2h_bytes;
    uint64_t tracked_objects, tracked_bytes, faults, uploads, allocs_seen, async_returns;
    uint64_t evictions, served, batches, max_batch, deferred, over_budget, resident_bytes;
    uint64_t serve_wait_ns, serve_run_ns, serve_launch_ns;
} stats_t;

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

static void one_call(worker_t *w, uint64_t *s, int *opout)
{
    const double u = unif(s);
    const double one = 1.0, zero = 0.0;
    const float onef = 1.0f, zerof = 0.0f;
    if (u < 0.6) {
        const int single = unif(s) < 0.5;
        const int m = logu(s, 32, 2048), n = logu(s, 32, 2048), k = logu(s, 32, 2048);
        const char ta = unif(s) < 0.5 ? 'N' : 'T', tb = unif(s) < 0.5 ? 'N' : 'T';
        const int ra = ta == 'N' ? m : k, ca = ta == 'N' ? k : m;
        const int rb = tb == 'N' ? k : n, cb = tb == 'N' ? n : k;
        const size_t ia = (size_t)(unif(s) * (GP - ra)) + (size_t)(unif(s) * (GP - ca)) * GP;
        const size_t ib = (size_t)(unif(s) * (GP - rb)) + (size_t)(unif(s) * (GP - cb)) * GP;
        const int ld = GP, ldc = m;
        if (single) {
            p_sgemm(&ta, &tb, &m, &n, &k, &onef, g_pools + ia, &ld, g_pools + ib, &ld, &zerof, w->cs,
                    &ldc);
            *opout = OP_SGEMM;
        } else {
            p_dgemm(&ta, &tb, &m, &n, &k, &one, g_poold + ia, &ld, g_poold + ib, &ld, &zero, w->cd,
                    &ldc);
            *opout = OP_DGEMM;
        }
    } else if (u < 0.9) {
        const int m = logu(s, 256, VP), n = logu(s, 256, VP);
        const char tr = unif(s) < 0.5 ? 'N' : 'T';
        const size_t ia = (size_t)(unif(s) * (VP - m)) + (size_t)(unif(s) * (VP - n)) * VP;
        const int ld = VP, inc = 1;
        const int lenx = tr == 'N' ? n : m;
        const size_t ix = (size_t)(unif(s) * (VP - lenx));
        p_dgemv(&tr, &m, &n, &one, g_poolv + ia, &ld, g_poolx + ix, &inc, &zero, w->y, &inc);
        *opout = OP_DGEMV;
    } else {
        const int m = logu(s, 64, TP), n = NRHS, ld = TP;
        const double alpha = 1024.0;
        p_dtrsm("L", "L", "N", "N", &m, &n, &alpha, g_poolt, &ld, w->bt, &ld);
        *opout = OP_DTRSM;
    }
    if (g_plain) p_sync();
}

static void *worker_main(void *arg)
{
    worker_t *w = arg;
    uint64_t s = g_seed * 1000003ull + (uint64_t)w->id;
    const int per = g_calls / g_threads;
    int dummy;
    for (int i = 0; i < g_warmup; ++i) one_call(w, &s, &dummy);
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
        one_call(w, &s, &w->op[i]);
        w->lat[i] = now() - t_sched;
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
    const char *lib = NULL, *opts = "";
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "--lib") && i + 1 < argc) lib = argv[++i];
        else if (!strcmp(argv[i], "--threads") && i + 1 < argc) g_threads = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--calls") && i + 1 < argc) g_calls = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--warmup") && i + 1 < argc) g_warmup = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--rate") && i + 1 < argc) g_rate = atof(argv[++i]);
        else if (!strcmp(argv[i], "--seed") && i + 1 < argc) g_seed = strtoull(argv[++i], 0, 10);
        else if (!strcmp(argv[i], "--plain")) g_plain = 1;
        else if (!strcmp(argv[i], "--options") && i + 1 < argc) opts = argv[++i];
    }
    if (!lib) {
        fprintf(stderr, "usage: serve_bench --lib libgpublas.so [--plain] [--options o] "
                        "[--rate calls/s] [--threads t] [--calls n]\n");
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
           "\"deferred\": %llu, \"evictions\": %llu, \"resident_bytes\": %llu, \"setup_s\": %.2f}\n",
           (unsigned long long)st.served, (unsigned long long)st.batches,
           st.batches ? (double)st.served / (double)st.batches : 0.0,
           (unsigned long long)st.max_batch, (unsigned long long)st.deferred,
           (unsigned long long)st.evictions, (unsigned long long)st.resident_bytes, t_setup);
    fflush(stdout);
    return 0;
}
