/*
 * chain_bench.c -- the memory-bound Level-1/2 chain as an ordinary BLAS
 * client: it allocates its operands with malloc and calls the Fortran BLAS
 * symbols, knowing nothing of the GPU.  Run it with
 *   LD_PRELOAD=libgpublas_preload.so GPUBLAS_OPTIONS="heuristic=objtrack:<f>"
 * and the allocations listed in <f> (its first four: A, x, y, z) become
 * HBM-resident tracked objects; run it with --blas <lib> [--prefix p]
 * [--threads T] and the same loop runs on that host BLAS instead (T threads
 * through OpenBLAS's <p>openblas_set_num_threads where the library has it).
 *
 * Per iteration (n = 8192 by default, FP64, entries uniform[-1,1]):
 *   y = A*x (dgemv 'N', alpha 1, beta 0); z = 0.3*y + z (daxpy);
 *   d = y.z (ddot); r = ||z|| (dnrm2); i = idamax(z)
 * Prints one JSON object with the wall time of the timed iterations.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static int cmp_double(const void *a, const void *b)
{
    const double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : x > y;
}

typedef void (*gemv_t)(const char *, const int *, const int *, const double *, const double *,
                       const int *, const double *, const int *, const double *, double *,
                       const int *);
typedef void (*axpy_t)(const int *, const double *, const double *, const int *, double *,
                       const int *);
typedef double (*dot_t)(const int *, const double *, const int *, const double *, const int *);
typedef double (*nrm2_t)(const int *, const double *, const int *);
typedef int (*iamax_t)(const int *, const double *, const int *);

static uint64_t splitmix_at(uint64_t seed, uint64_t i)
{
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void fill(double *p, size_t n, uint64_t seed)
{
    for (size_t i = 0; i < n; ++i) p[i] = 2.0 * ((double)(splitmix_at(seed, i) >> 11) * 0x1.0p-53) - 1.0;
}

static double now(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC_RAW, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

static void *sym(void *h, const char *prefix, const char *name)
{
    char nm[128];
    snprintf(nm, sizeof nm, "%s%s", prefix, name);
    void *f = dlsym(h, nm);
    if (!f) {
        fprintf(stderr, "chain_bench: no symbol %s\n", nm);
        exit(2);
    }
    return f;
}

int main(int argc, char **argv)
{
    int n = 8192, iters = 200, warmup = 5;
    uint64_t seed = 20240601;
    const char *blas = NULL, *prefix = "", *direct = NULL;
    int threads = 0; /* --blas: OpenBLAS threads (0: its default) */
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "--n") && i + 1 < argc) n = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--iters") && i + 1 < argc) iters = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--warmup") && i + 1 < argc) warmup = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--seed") && i + 1 < argc) seed = strtoull(argv[++i], 0, 10);
        else if (!strcmp(argv[i], "--blas") && i + 1 < argc) blas = argv[++i];
        else if (!strcmp(argv[i], "--prefix") && i + 1 < argc) prefix = argv[++i];
        else if (!strcmp(argv[i], "--lib") && i + 1 < argc) direct = argv[++i];
        else if (!strcmp(argv[i], "--threads") && i + 1 < argc) threads = atoi(argv[++i]);
    }
    void *(*alloc)(size_t) = malloc;
    void (*release)(void *) = free;
    void *dh = NULL;
    if (direct) {
        dh = dlopen(direct, RTLD_NOW | RTLD_GLOBAL);
        if (!dh) {
            fprintf(stderr, "chain_bench: %s\n", dlerror());
            return 2;
        }
        alloc = (void *(*)(size_t))dlsym(dh, "gpublas_malloc");
        release = (void (*)(void *))dlsym(dh, "gpublas_free");
        if (!alloc || !release) {
            fprintf(stderr, "chain_bench: %s lacks gpublas_malloc\n", direct);
            return 2;
        }
    }
    /* The operands are this program's first four allocations. */
    const size_t nn = (size_t)n * n;
    double *A = alloc(nn * sizeof(double));
    double *x = alloc((size_t)n * sizeof(double));
    double *y = alloc((size_t)n * sizeof(double));
    double *z = alloc((size_t)n * sizeof(double));
    if (!A || !x || !y || !z) {
        fprintf(stderr, "chain_bench: out of memory\n");
        return 2;
    }
    void *h = dh ? dh : RTLD_DEFAULT;
    if (blas) {
        h = dlopen(blas, RTLD_NOW | RTLD_LOCAL);
        if (!h) {
            fprintf(stderr, "chain_bench: %s\n", dlerror());
            return 2;
        }
    }
    int blas_threads = -1;
    if (blas) {
        /* OpenBLAS's thread controls, when the library has them */
        char nm[128];
        snprintf(nm, sizeof nm, "%sopenblas_set_num_threads", prefix);
        void (*set_threads)(int) = (void (*)(int))dlsym(h, nm);
        snprintf(nm, sizeof nm, "%sopenblas_get_num_threads", prefix);
        int (*get_threads)(void) = (int (*)(void))dlsym(h, nm);
        if (set_threads && threads > 0) set_threads(threads);
        if (get_threads) blas_threads = get_threads();
    }
    gemv_t gemv = (gemv_t)sym(h, prefix, "dgemv_");
    axpy_t axpy = (axpy_t)sym(h, prefix, "daxpy_");
    dot_t dot = (dot_t)sym(h, prefix, "ddot_");
    nrm2_t nrm2 = (nrm2_t)sym(h, prefix, "dnrm2_");
    iamax_t iamax = (iamax_t)sym(h, prefix, "idamax_");
    void *gh = dh ? dh : RTLD_DEFAULT;
    int (*is_tracked)(const void *) = (int (*)(const void *))dlsym(gh, "gpublas_is_tracked");
    int (*ktime)(const char *, double *) = (int (*)(const char *, double *))dlsym(gh, "gpublas_kernel_time");
    void (*prof_reset)(void) = (void (*)(void))dlsym(gh, "gpublas_dev_prof_reset");

    fill(A, nn, seed);
    fill(x, n, seed + 1);
    memset(y, 0, (size_t)n * sizeof(double));
    fill(z, n, seed + 3);

    const double one = 1.0, zero = 0.0, a3 = 0.3;
    const int inc = 1;
    double d = 0, r = 0;
    int im = 0;
    double t0 = 0, tp = 0;
    /* per-iteration times: every iteration ends in idamax, whose result the
     * caller holds, so the iteration's work is complete at its end */
    double *it_s = iters > 0 ? malloc((size_t)iters * sizeof(double)) : NULL;
    for (int it = 0; it < warmup + iters; ++it) {
        if (it == warmup) {
            if (prof_reset) prof_reset();
            t0 = tp = now();
        }
        gemv("N", &n, &n, &one, A, &n, x, &inc, &zero, y, &inc);
        axpy(&n, &a3, y, &inc, z, &inc);
        d = dot(&n, y, &inc, z, &inc);
        r = nrm2(&n, z, &inc);
        im = iamax(&n, z, &inc);
        if (it >= warmup && it_s) {
            const double tn = now();
            it_s[it - warmup] = tn - tp;
            tp = tn;
        }
    }
    const double secs = now() - t0;
    const double bytes = ((double)nn + 9.0 * n) * 8.0 * iters;
    int tracked = -1;
    if (is_tracked) tracked = is_tracked(A) + is_tracked(x) + is_tracked(y) + is_tracked(z);
    printf("{\"n\": %d, \"iters\": %d, \"warmup\": %d, \"seconds\": %.6f, \"ms_per_iter\": %.6f, "
           "\"GBps\": %.3f, \"ddot\": %.17g, \"dnrm2\": %.17g, \"idamax\": %d, \"tracked\": %d",
           n, iters, warmup, secs, 1e3 * secs / iters, bytes / secs / 1e9, d, r, im, tracked);
    if (it_s) {
        qsort(it_s, (size_t)iters, sizeof(double), cmp_double);
        const double p10 = it_s[(iters - 1) / 10], p50 = it_s[(iters - 1) / 2],
                     p90 = it_s[(iters - 1) * 9 / 10];
        printf(", \"iter_ms_p10\": %.6f, \"iter_ms_p50\": %.6f, \"iter_ms_p90\": %.6f", 1e3 * p10,
               1e3 * p50, 1e3 * p90);
        free(it_s);
    }
    if (blas) printf(", \"blas_threads\": %d", blas_threads);
    if (ktime) {
        const char *ks[] = {"dgemv_n", "daxpy", "ddot", "dnrm2", "idamax"};
        for (int i = 0; i < 5; ++i) {
            double ms = 0;
            int c = ktime(ks[i], &ms);
            printf(", \"k_%s_ms\": %.6f, \"k_%s_count\": %d", ks[i], ms, ks[i], c);
        }
    }
    printf("}\n");
    fflush(stdout);
    /* y and z are read back through the host view: exercises coherence */
    double ysum = 0;
    for (int i = 0; i < n; ++i) ysum += y[i] + z[i];
    fprintf(stderr, "chain_bench: checksum %.17g\n", ysum);
    release(A);
    release(x);
    release(y);
    release(z);
    return 0;
}
