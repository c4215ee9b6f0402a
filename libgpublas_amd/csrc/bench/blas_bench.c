/*
 * blas_bench.c -- times one Level-3 (or Level-1 calibration) call shape
 * through the Fortran-ABI entry points of libgpublas on tracked operands.
 *
 *   blas_bench --lib libgpublas.so --op gemm --prec d --m M --n N --k K --ta N --tb N
 *                                    (--prec s, z or c: SGEMM, ZGEMM, CGEMM; complex flops
 *                                     counted as 8mnk real flops; syrk, syr2k, trsm, trmm,
 *                                     symm take z / c as well, plus herk, her2k, hemm)
 *   blas_bench --lib ... --op syrk --n N --k K --uplo L --trans N
 *   blas_bench --lib ... --op trsm --m M --n N --side L --uplo L --ta N --diag N
 *   blas_bench --lib ... --op symm --m M --n N --side L --uplo L
 *   blas_bench --lib ... --op syr2k --n N --k K --uplo L --trans N
 *   blas_bench --lib ... --op trmm --m M --n N --side L --uplo L --ta N --diag N
 *   blas_bench --lib ... --op gemv --m M --n N --ta N (DGEMV, N or T)
 *   blas_bench --lib ... --op symv --n N --uplo L   (DSYMV, bytes = the triangle once)
 *   blas_bench --lib ... --op syr --n N --uplo L    (DSYR, the triangle read and written)
 *   blas_bench --lib ... --op trmv --n N --uplo L --ta N --diag N  (DTRMV, the triangle once)
 *   blas_bench --lib ... --op trsv --n N --uplo L --ta N --diag N  (DTRSV, the triangle once;
 *                                                   diagonally dominant triangle)
 *   blas_bench --lib ... --op ger --m M --n N        (DGER, A read and written once)
 *   blas_bench --lib ... --op spmv --n N --uplo L   (DSPMV on packed storage)
 *   blas_bench --lib ... --op tpmv|tpsv --n N --uplo L --ta N --diag N   (DTPMV / DTPSV, packed)
 *   blas_bench --lib ... --op gbmv --m M --n N --k KL=KU --ta N   (DGBMV, the band once)
 *   blas_bench --lib ... --op sbmv|tbmv|tbsv --n N --k K --uplo L [--ta N --diag N]
 *              [--dominant]  (tbsv: diagonally dominant band)
 *   blas_bench --lib ... --op axpy --n N           (HBM traffic calibration)
 *   blas_bench --lib ... --op link --n BYTES        (host-link copies of pageable
 *                                                    memory: plain vs staged)
 *   common: --reps R --warmup W --seed S
 *
 * Operands come from gpublas_malloc and are filled on the device with the
 * library's SplitMix64 generator (uniform[-1,1]); the TRSM triangle gets
 * diag = n + uniform[0,1] on the host.  Prints one JSON object: wall time per
 * call (calls on resident operands may return before their kernels finish,
 * so the timed loop is closed by a stream synchronisation), the kernel times
 * from the library's HIP-event profiler (GPUBLAS_OPTIONS=profile), and the
 * rate.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC_RAW, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

static void *lib;
static void *sym(const char *n)
{
    void *f = dlsym(lib, n);
    if (!f) {
        fprintf(stderr, "blas_bench: missing %s\n", n);
        exit(2);
    }
    return f;
}

typedef void *(*malloc_t)(size_t);
typedef void *(*devptr_t)(const void *);
typedef int (*fill_t)(void *, uint64_t, int64_t, void *);
typedef int (*fill32_t)(void *, uint64_t, int64_t, void *);

static malloc_t g_malloc;
static devptr_t g_devptr;
static fill_t g_fill;
static fill32_t g_fill32;
static int (*g_sync)(void *);

static void *operand(size_t elems, int prec, uint64_t seed)
{
    if (prec == 'z' || prec == 'c') { /* complex: (re, im) pairs of the real fill */
        elems *= 2;
        prec = prec == 'z' ? 'd' : 's';
    }
    size_t es = prec == 'd' ? 8 : 4;
    void *h = g_malloc(elems * es);
    if (!h) {
        fprintf(stderr, "blas_bench: gpublas_malloc(%zu) failed\n", elems * es);
        exit(2);
    }
    void *d = g_devptr(h);
    if (prec == 'd') g_fill(NULL, seed, (int64_t)elems, d);
    else g_fill32(NULL, seed, (int64_t)elems, d);
    g_sync(NULL);
    return h;
}

/* "<prec><op>_": the Fortran entry of a routine in the chosen precision */
static const char *fname(int prec, const char *op)
{
    static char buf[8][16];
    static int i;
    i = (i + 1) & 7;
    snprintf(buf[i], sizeof buf[i], "%c%s_", prec, op);
    return buf[i];
}

int main(int argc, char **argv)
{
    const char *path = NULL, *op = "gemm";
    int prec = 'd', m = 1024, n = 1024, k = 1024, reps = 5, warmup = 1;
    char ta = 'N', tb = 'N', uplo = 'L', trans = 'N', side = 'L', diag = 'N';
    uint64_t seed = 7;
    int dominant = 0; /* tbsv: diagonally dominant band (see below) */
    for (int i = 1; i < argc; ++i) {
        const char *a = argv[i], *v = i + 1 < argc ? argv[i + 1] : "";
        if (!strcmp(a, "--lib")) path = v, ++i;
        else if (!strcmp(a, "--op")) op = v, ++i;
        else if (!strcmp(a, "--prec")) prec = v[0], ++i;
        else if (!strcmp(a, "--m")) m = atoi(v), ++i;
        else if (!strcmp(a, "--n")) n = atoi(v), ++i;
        else if (!strcmp(a, "--k")) k = atoi(v), ++i;
        else if (!strcmp(a, "--ta")) ta = v[0], ++i;
        else if (!strcmp(a, "--tb")) tb = v[0], ++i;
        else if (!strcmp(a, "--uplo")) uplo = v[0], ++i;
        else if (!strcmp(a, "--trans")) trans = v[0], ++i;
        else if (!strcmp(a, "--side")) side = v[0], ++i;
        else if (!strcmp(a, "--diag")) diag = v[0], ++i;
        else if (!strcmp(a, "--reps")) reps = atoi(v), ++i;
        else if (!strcmp(a, "--warmup")) warmup = atoi(v), ++i;
        else if (!strcmp(a, "--seed")) seed = strtoull(v, 0, 10), ++i;
        else if (!strcmp(a, "--dominant")) dominant = 1;
    }
    if (!path) {
        fprintf(stderr, "blas_bench: --lib required\n");
        return 2;
    }
    lib = dlopen(path, RTLD_NOW | RTLD_GLOBAL);
    if (!lib) {
        fprintf(stderr, "blas_bench: %s\n", dlerror());
        return 2;
    }
    g_malloc = (malloc_t)sym("gpublas_malloc");
    g_devptr = (devptr_t)sym("gpublas_device_ptr");
    g_fill = (fill_t)sym("gpublas_dev_fill_uniform");
    g_fill32 = (fill32_t)sym("gpublas_dev_fill_uniform_f32");
    g_sync = (int (*)(void *))sym("gpublas_dev_sync");
    void (*set_cutoff)(int) = (void (*)(int))sym("gpublas_set_cutoff");
    int (*ktime)(const char *, double *) = (int (*)(const char *, double *))sym("gpublas_kernel_time");
    int (*dump)(char *, size_t) = (int (*)(char *, size_t))sym("gpublas_dev_prof_dump");
    void (*prof_reset)(void) = (void (*)(void))sym("gpublas_dev_prof_reset");
    set_cutoff(0);
    (void)ktime;

    if (!strcmp(op, "link")) {
        /* best-of-reps time of one contiguous copy each way, plain pageable
         * hipMemcpyAsync against the library's staged copy; each timed copy
         * ends with a stream synchronisation */
        typedef int (*cp_t)(void *, const void *, size_t, void *);
        typedef int (*cps_t)(void *, size_t, const void *, size_t, size_t, size_t, size_t, void *);
        cp_t h2d = (cp_t)sym("gpublas_dev_h2d"), d2h = (cp_t)sym("gpublas_dev_d2h");
        cps_t h2ds = (cps_t)sym("gpublas_dev_h2d_staged"), d2hs = (cps_t)sym("gpublas_dev_d2h_staged");
        void *(*dalloc)(size_t) = (void *(*)(size_t))sym("gpublas_dev_alloc");
        const size_t nb = (size_t)n;
        char *h = malloc(nb), *h2 = malloc(nb);
        void *d = dalloc(nb);
        if (!h || !h2 || !d) {
            fprintf(stderr, "blas_bench: cannot allocate %zu bytes\n", nb);
            return 2;
        }
        memset(h, 1, nb);
        memset(h2, 2, nb);
        double t[4] = {1e30, 1e30, 1e30, 1e30};
        for (int r = 0; r < warmup + reps; ++r) {
            double t0 = now(), dt;
            h2d(d, h, nb, NULL), g_sync(NULL), dt = now() - t0;
            if (r >= warmup && dt < t[0]) t[0] = dt;
            t0 = now(), h2ds(d, nb, h, nb, nb, 1, 1, NULL), g_sync(NULL), dt = now() - t0;
            if (r >= warmup && dt < t[1]) t[1] = dt;
            t0 = now(), d2h(h2, d, nb, NULL), g_sync(NULL), dt = now() - t0;
            if (r >= warmup && dt < t[2]) t[2] = dt;
            t0 = now(), d2hs(h2, nb, d, nb, nb, 1, 1, NULL), g_sync(NULL), dt = now() - t0;
            if (r >= warmup && dt < t[3]) t[3] = dt;
        }
        if (memcmp(h, h2, nb) != 0) {
            fprintf(stderr, "blas_bench: link copy round trip differs\n");
            return 1;
        }
        printf("{\"op\": \"link\", \"bytes\": %zu, \"upload_pageable_us\": %.1f, "
               "\"upload_staged_us\": %.1f, \"download_pageable_us\": %.1f, "
               "\"download_staged_us\": %.1f, \"upload_pageable_gbs\": %.1f, "
               "\"upload_staged_gbs\": %.1f, \"download_pageable_gbs\": %.1f, "
               "\"download_staged_gbs\": %.1f}\n",
               nb, t[0] * 1e6, t[1] * 1e6, t[2] * 1e6, t[3] * 1e6, nb / t[0] / 1e9,
               nb / t[1] / 1e9, nb / t[2] / 1e9, nb / t[3] / 1e9);
        return 0;
    }

    double flops = 0, bytes = 0;
    void *A = NULL, *B = NULL, *C = NULL;
    double alpha_d[2] = {1.0, 0.0}, beta_d[2] = {0.5, 0.0};
    float alpha_s[2] = {1.0f, 0.0f}, beta_s[2] = {0.5f, 0.0f};
    const int dbl = prec == 'd' || prec == 'z';
    void *alpha = dbl ? (void *)alpha_d : (void *)alpha_s;
    void *beta = dbl ? (void *)beta_d : (void *)beta_s;
    int lda = 0, ldb = 0, ldc = 0;
    typedef void (*gemm_t)(const char *, const char *, const int *, const int *, const int *,
                           const void *, const void *, const int *, const void *, const int *,
                           const void *, void *, const int *);
    typedef void (*syrk_t)(const char *, const char *, const int *, const int *, const void *,
                           const void *, const int *, const void *, void *, const int *);
    typedef void (*trsm_t)(const char *, const char *, const char *, const char *, const int *,
                           const int *, const void *, const void *, const int *, void *,
                           const int *);
    typedef void (*axpy_t)(const int *, const double *, const double *, const int *, double *,
                           const int *);
    typedef void (*symm_t)(const char *, const char *, const int *, const int *, const void *,
                           const void *, const int *, const void *, const int *, const void *,
                           void *, const int *);
    gemm_t gemm = NULL;
    syrk_t syrk = NULL;
    trsm_t trsm = NULL;
    axpy_t axpy = NULL;
    symm_t symm = NULL;
    typedef void (*syr2k_t)(const char *, const char *, const int *, const int *, const void *,
                            const void *, const int *, const void *, const int *, const void *,
                            void *, const int *);
    syr2k_t syr2k = NULL;
    trsm_t trmm = NULL; /* same argument list as TRSM */
    typedef void (*symv_t)(const char *, const int *, const double *, const double *, const int *,
                           const double *, const int *, const double *, double *, const int *);
    typedef void (*syr_t)(const char *, const int *, const double *, const double *, const int *,
                          double *, const int *);
    symv_t symv = NULL;
    typedef void (*gemv_t)(const char *, const int *, const int *, const double *, const double *,
                           const int *, const double *, const int *, const double *, double *,
                           const int *);
    gemv_t gemv = NULL;
    syr_t syr = NULL;
    typedef void (*trmv_t)(const char *, const char *, const char *, const int *, const double *,
                           const int *, double *, const int *);
    trmv_t trmv = NULL;
    trmv_t trsv = NULL; /* same argument list as DTRMV */
    typedef void (*ger_t)(const int *, const int *, const double *, const double *, const int *,
                          const double *, const int *, double *, const int *);
    ger_t ger = NULL;
    typedef void (*spmv_t)(const char *, const int *, const double *, const double *,
                           const double *, const int *, const double *, double *, const int *);
    spmv_t spmv = NULL;
    typedef void (*tpmv_t)(const char *, const char *, const char *, const int *, const double *,
                           double *, const int *);
    tpmv_t tpmv = NULL;
    typedef void (*gbmv_t)(const char *, const int *, const int *, const int *, const int *,
                           const double *, const double *, const int *, const double *,
                           const int *, const double *, double *, const int *);
    gbmv_t gbmv = NULL;
    typedef void (*sbmv_t)(const char *, const int *, const int *, const double *, const double *,
                           const int *, const double *, const int *, const double *, double *,
                           const int *);
    sbmv_t sbmv = NULL;
    typedef void (*tbmv_t)(const char *, const char *, const char *, const int *, const int *,
                           const double *, const int *, double *, const int *);
    tbmv_t tbmv = NULL;
    int bw = 0; /* band array rows */
    const double cx = (prec == 'z' || prec == 'c') ? 4.0 : 1.0; /* real flops per FMA over 2 */
    if (!strcmp(op, "gemm")) {
        gemm = (gemm_t)sym(prec == 'd' ? "dgemm_" : prec == 's' ? "sgemm_" : prec == 'z' ? "zgemm_" : "cgemm_");
        lda = (ta == 'N' || ta == 'n') ? m : k;
        ldb = (tb == 'N' || tb == 'n') ? k : n;
        ldc = m;
        A = operand((size_t)lda * ((ta == 'N' || ta == 'n') ? k : m), prec, seed);
        B = operand((size_t)ldb * ((tb == 'N' || tb == 'n') ? n : k), prec, seed + 1);
        C = operand((size_t)m * n, prec, seed + 2);
        flops = (prec == 'z' || prec == 'c' ? 8.0 : 2.0) * m * n * k; /* real flops */
    } else if (!strcmp(op, "syrk") || !strcmp(op, "herk")) {
        syrk = (syrk_t)sym(fname(prec, op));
        lda = (trans == 'N' || trans == 'n') ? n : k;
        ldc = n;
        A = operand((size_t)lda * ((trans == 'N' || trans == 'n') ? k : n), prec, seed);
        C = operand((size_t)n * n, prec, seed + 2);
        flops = cx * n * (n + 1) * k;
    } else if (!strcmp(op, "trsm")) {
        trsm = (trsm_t)sym(fname(prec, "trsm"));
        const int na = (side == 'L' || side == 'l') ? m : n;
        lda = na;
        ldb = m;
        A = operand((size_t)na * na, prec, seed);
        B = operand((size_t)m * n, prec, seed + 1);
        /* diagonally dominant: diag = n + uniform[0,1] (host touch of the diagonal) */
        uint64_t s = seed + 5;
        for (int i = 0; i < na; ++i) {
            s = s * 6364136223846793005ull + 1442695040888963407ull;
            double u = (double)(s >> 11) * 0x1.0p-53;
            const size_t d = (size_t)i * na + i, e = (prec == 'z' || prec == 'c') ? 2 * d : d;
            if (prec == 'd' || prec == 'z') ((double *)A)[e] = na + u;
            else ((float *)A)[e] = (float)(na + u);
        }
        flops = cx * ((side == 'L' || side == 'l') ? (double)m * m * n : (double)n * n * m);
    } else if (!strcmp(op, "syr2k") || !strcmp(op, "her2k")) {
        syr2k = (syr2k_t)sym(fname(prec, op));
        lda = ldb = (trans == 'N' || trans == 'n') ? n : k;
        ldc = n;
        A = operand((size_t)lda * ((trans == 'N' || trans == 'n') ? k : n), prec, seed);
        B = operand((size_t)ldb * ((trans == 'N' || trans == 'n') ? k : n), prec, seed + 1);
        C = operand((size_t)n * n, prec, seed + 2);
        flops = cx * 2.0 * n * (n + 1) * k;
    } else if (!strcmp(op, "trmm")) {
        trmm = (trsm_t)sym(fname(prec, "trmm"));
        const int na = (side == 'L' || side == 'l') ? m : n;
        lda = na;
        ldb = m;
        A = operand((size_t)na * na, prec, seed);
        B = operand((size_t)m * n, prec, seed + 1);
        flops = cx * ((side == 'L' || side == 'l') ? (double)m * m * n : (double)n * n * m);
    } else if (!strcmp(op, "symm") || !strcmp(op, "hemm")) {
        symm = (symm_t)sym(fname(prec, op));
        const int na = (side == 'L' || side == 'l') ? m : n;
        lda = na;
        ldb = ldc = m;
        A = operand((size_t)na * na, prec, seed);
        B = operand((size_t)m * n, prec, seed + 1);
        C = operand((size_t)m * n, prec, seed + 2);
        flops = cx * 2.0 * m * n * na;
    } else if (!strcmp(op, "gemv")) {
        gemv = (gemv_t)sym("dgemv_");
        A = operand((size_t)m * n, 'd', seed);
        B = operand((size_t)(ta == 'N' ? n : m), 'd', seed + 1);
        C = operand((size_t)(ta == 'N' ? m : n), 'd', seed + 2);
        bytes = 8.0 * ((double)m * n + m + n);
        flops = 2.0 * m * n;
    } else if (!strcmp(op, "symv")) {
        symv = (symv_t)sym("dsymv_");
        A = operand((size_t)n * n, 'd', seed);
        B = operand((size_t)n, 'd', seed + 1);
        C = operand((size_t)n, 'd', seed + 2);
        bytes = 8.0 * ((double)n * (n + 1) / 2 + 3.0 * n); /* triangle once, x, y in/out */
        flops = 2.0 * n * n;
    } else if (!strcmp(op, "syr")) {
        syr = (syr_t)sym("dsyr_");
        A = operand((size_t)n * n, 'd', seed);
        B = operand((size_t)n, 'd', seed + 1);
        bytes = 8.0 * ((double)n * (n + 1) + n); /* triangle read and written, x */
        flops = (double)n * n;
    } else if (!strcmp(op, "trmv")) {
        trmv = (trmv_t)sym("dtrmv_");
        A = operand((size_t)n * n, 'd', seed);
        B = operand((size_t)n, 'd', seed + 1);
        bytes = 8.0 * ((double)n * (n + 1) / 2 + 2.0 * n); /* triangle once, x in/out */
        flops = (double)n * n;
    } else if (!strcmp(op, "trsv")) {
        trsv = (trmv_t)sym("dtrsv_");
        A = operand((size_t)n * n, 'd', seed);
        B = operand((size_t)n, 'd', seed + 1);
        /* diagonally dominant: diag = n + uniform[0,1] (host touch of the diagonal) */
        uint64_t s = seed + 5;
        for (int i = 0; i < n; ++i) {
            s = s * 6364136223846793005ull + 1442695040888963407ull;
            ((double *)A)[(size_t)i * n + i] = n + (double)(s >> 11) * 0x1.0p-53;
        }
        bytes = 8.0 * ((double)n * (n + 1) / 2 + 2.0 * n); /* triangle once, x in/out */
        flops = (double)n * n;
    } else if (!strcmp(op, "ger")) {
        ger = (ger_t)sym("dger_");
        A = operand((size_t)m * n, 'd', seed);
        B = operand((size_t)m, 'd', seed + 1);
        C = operand((size_t)n, 'd', seed + 2);
        bytes = 8.0 * (2.0 * m * n + m + n); /* A read and written, x, y */
        flops = 2.0 * m * n;
    } else if (!strcmp(op, "spmv")) {
        spmv = (spmv_t)sym("dspmv_");
        A = operand((size_t)n * (n + 1) / 2, 'd', seed);
        B = operand((size_t)n, 'd', seed + 1);
        C = operand((size_t)n, 'd', seed + 2);
        bytes = 8.0 * ((double)n * (n + 1) / 2 + 3.0 * n); /* packed triangle once, x, y */
        flops = 2.0 * n * n;
    } else if (!strcmp(op, "tpmv") || !strcmp(op, "tpsv")) {
        tpmv = (tpmv_t)sym(op[2] == 'm' ? "dtpmv_" : "dtpsv_");
        A = operand((size_t)n * (n + 1) / 2, 'd', seed);
        B = operand((size_t)n, 'd', seed + 1);
        if (op[2] == 's') { /* as for trsv: diagonal n + uniform[0,1] (packed positions) */
            uint64_t s = seed + 5;
            const int lo = uplo == 'L' || uplo == 'l';
            for (size_t j = 0; j < (size_t)n; ++j) {
                s = s * 6364136223846793005ull + 1442695040888963407ull;
                const size_t d = lo ? j * (2 * (size_t)n - j + 1) / 2 : j * (j + 1) / 2 + j;
                ((double *)A)[d] = n + (double)(s >> 11) * 0x1.0p-53;
            }
        }
        bytes = 8.0 * ((double)n * (n + 1) / 2 + 2.0 * n); /* packed triangle once, x in and out */
        flops = 1.0 * n * n;
    } else if (!strcmp(op, "gbmv")) {
        gbmv = (gbmv_t)sym("dgbmv_");
        bw = 2 * k + 1;
        A = operand((size_t)bw * n, 'd', seed);
        B = operand((size_t)(ta == 'N' ? n : m), 'd', seed + 1);
        C = operand((size_t)(ta == 'N' ? m : n), 'd', seed + 2);
        bytes = 8.0 * ((double)bw * n + m + 2.0 * (ta == 'N' ? m : n)); /* band, x, y in/out */
        flops = 2.0 * bw * n;
    } else if (!strcmp(op, "sbmv") || !strcmp(op, "tbmv") || !strcmp(op, "tbsv")) {
        bw = k + 1;
        A = operand((size_t)bw * n, 'd', seed);
        B = operand((size_t)n, 'd', seed + 1);
        if (op[0] == 's') {
            sbmv = (sbmv_t)sym("dsbmv_");
            C = operand((size_t)n, 'd', seed + 2);
            bytes = 8.0 * ((double)bw * n + 3.0 * n); /* the stored band once, x, y in/out */
            flops = 4.0 * bw * n;
        } else {
            tbmv = (tbmv_t)sym(op[2] == 'm' ? "dtbmv_" : "dtbsv_");
            if (dominant) {
                /* off-diagonal entries u / (2 (k + 1)), diagonal 2 + u / 2 (u
                 * the uniform fill in [-1, 1]): host writes, so the first call
                 * uploads the band again */
                double *ab = (double *)A;
                const int dr = uplo == 'U' || uplo == 'u' ? k : 0;
                for (size_t j = 0; j < (size_t)n; ++j)
                    for (int r = 0; r < bw; ++r)
                        ab[j * bw + r] = r == dr ? 2.0 + 0.5 * ab[j * bw + r]
                                                 : ab[j * bw + r] * (0.5 / bw);
            }
            bytes = 8.0 * ((double)bw * n + 2.0 * n);
            flops = 2.0 * bw * n;
        }
    } else if (!strcmp(op, "axpy")) {
        axpy = (axpy_t)sym("daxpy_");
        A = operand((size_t)n, 'd', seed);
        B = operand((size_t)n, 'd', seed + 1);
        bytes = 24.0 * n;
    } else {
        fprintf(stderr, "blas_bench: unknown op %s\n", op);
        return 2;
    }
    double t0 = 0;
    for (int r = 0; r < warmup + reps; ++r) {
        if (r == warmup) {
            g_sync(NULL); /* calls on resident operands may return before their kernels */
            prof_reset();
            t0 = now();
        }
        if (gemm) gemm(&ta, &tb, &m, &n, &k, alpha, A, &lda, B, &ldb, beta, C, &ldc);
        else if (syrk) syrk(&uplo, &trans, &n, &k, alpha, A, &lda, beta, C, &ldc);
        else if (trsm) trsm(&side, &uplo, &ta, &diag, &m, &n, alpha, A, &lda, B, &ldb);
        else if (symm) symm(&side, &uplo, &m, &n, alpha, A, &lda, B, &ldb, beta, C, &ldc);
        else if (syr2k) syr2k(&uplo, &trans, &n, &k, alpha, A, &lda, B, &ldb, beta, C, &ldc);
        else if (trmm) trmm(&side, &uplo, &ta, &diag, &m, &n, alpha, A, &lda, B, &ldb);
        else if (gemv) {
            const double a = 1.0, b = 0.5;
            const int one = 1;
            gemv(&ta, &m, &n, &a, (const double *)A, &m, (const double *)B, &one, &b, (double *)C,
                 &one);
        } else if (symv) {
            const double a = 1.0, b = 0.5;
            const int one = 1;
            symv(&uplo, &n, &a, (const double *)A, &n, (const double *)B, &one, &b, (double *)C,
                 &one);
        } else if (trmv) {
            const int one = 1;
            trmv(&uplo, &ta, &diag, &n, (const double *)A, &n, (double *)B, &one);
        } else if (trsv) {
            const int one = 1;
            trsv(&uplo, &ta, &diag, &n, (const double *)A, &n, (double *)B, &one);
        } else if (ger) {
            const double a = 1e-3;
            const int one = 1;
            ger(&m, &n, &a, (const double *)B, &one, (const double *)C, &one, (double *)A, &m);
        } else if (spmv) {
            const double a = 1.0, b = 0.5;
            const int one = 1;
            spmv(&uplo, &n, &a, (const double *)A, (const double *)B, &one, &b, (double *)C, &one);
        } else if (tpmv) {
            const int one = 1;
            tpmv(&uplo, &ta, &diag, &n, (const double *)A, (double *)B, &one);
        } else if (gbmv) {
            const double a = 1.0, b = 0.5;
            const int one = 1;
            gbmv(&ta, &m, &n, &k, &k, &a, (const double *)A, &bw, (const double *)B, &one, &b,
                 (double *)C, &one);
        } else if (sbmv) {
            const double a = 1.0, b = 0.5;
            const int one = 1;
            sbmv(&uplo, &n, &k, &a, (const double *)A, &bw, (const double *)B, &one, &b,
                 (double *)C, &one);
        } else if (tbmv) {
            const int one = 1;
            tbmv(&uplo, &ta, &diag, &n, &k, (const double *)A, &bw, (double *)B, &one);
        } else if (syr) {
            const double a = 1e-3;
            const int one = 1;
            syr(&uplo, &n, &a, (const double *)B, &one, (double *)A, &n);
        } else {
            const double a = 0.5;
            const int one = 1;
            axpy(&n, &a, (const double *)A, &one, (double *)B, &one);
        }
    }
    g_sync(NULL);
    const double wall = (now() - t0) / reps;
    char buf[4096];
    buf[0] = 0;
    int nb = dump(buf, sizeof buf);
    printf("{\"op\": \"%s\", \"prec\": \"%c\", \"m\": %d, \"n\": %d, \"k\": %d, \"ta\": \"%c\", "
           "\"tb\": \"%c\", \"uplo\": \"%c\", \"trans\": \"%c\", \"side\": \"%c\", \"reps\": %d, "
           "\"wall_ms\": %.6f",
           op, prec, m, n, k, ta, tb, uplo, trans, side, reps, wall * 1e3);
    if (flops > 0) printf(", \"flops\": %.6e, \"wall_tflops\": %.4f", flops, flops / wall / 1e12);
    if (bytes > 0) printf(", \"bytes\": %.6e, \"wall_GBps\": %.2f", bytes, bytes / wall / 1e9);
    /* kernels: "name count total_ms" lines */
    printf(", \"kernels\": {");
    int first = 1;
    for (char *line = strtok(nb > 0 ? buf : NULL, "\n"); line; line = strtok(NULL, "\n")) {
        char nm[128];
        long cnt;
        double ms;
        if (sscanf(line, "%127s %ld %lf", nm, &cnt, &ms) == 3 && cnt > 0) {
            printf("%s\"%s\": [%ld, %.6f]", first ? "" : ", ", nm, cnt, ms);
            first = 0;
        }
    }
    printf("}}\n");
    return 0;
}
