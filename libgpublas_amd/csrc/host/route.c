/*
 * route.c -- host or GPU: where a call whose operands are not resident runs.
 *
 * The reference sends a GEMM to the host BLAS when a leading dimension is
 * below 512 and every other routine to the GPU (blas_level3/gemm.cc:129-141);
 * its TODO.md:7 asks for the CPU BLAS to be chosen dynamically where that is
 * faster, and data/runtimes.txt:1-3 records why (a 32^3 DGEMM ran 8x slower
 * through the GPU than on the host).  Three modes:
 *
 *   cutoff=<n> (n > 0, default 512)  the reference's fixed rule: GEMM leading
 *                                    dimensions, and Level-1/2 lengths, below
 *                                    n go to the host;
 *   cutoff=0                         everything on the GPU;
 *   cutoff=auto                      a measured crossover table: a call runs on
 *                                    the GPU when its flops reach the
 *                                    routine's min_flops AND its flops per
 *                                    operand byte reach min_intensity.
 *
 * In every mode a call with at least one tracked (device-resident) operand
 * runs on the GPU, where its data already is.
 *
 * The two thresholds come from gpublas_tune_crossover(): for each routine it
 * times square problems of growing size on the host BLAS and on the device
 * path with operands that are not resident (so the device time includes the
 * host-link copies and the synchronisation) and takes the smallest size from
 * which the GPU is faster at every larger size measured (each side timed in
 * a phase of its own).  min_flops is that
 * size's flop count (the fixed cost of a device call) and min_intensity its
 * flops per byte copied (the link cost: a long thin product with the same
 * flops but far more bytes stays on the host).  A routine on which the GPU
 * never won gets +inf.  The table can be written to a text file and loaded
 * with xover=<file>; without one the built-in table below is used.
 */
#include "internal.h"
#include <dlfcn.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

GB_TLS int gb_route_gpu = 0;

static const char *const k_names[GB_RK_COUNT] = {
    [GB_RK_DGEMM] = "dgemm",   [GB_RK_SGEMM] = "sgemm",   [GB_RK_DSYRK] = "dsyrk",
    [GB_RK_SSYRK] = "ssyrk",   [GB_RK_DSYR2K] = "dsyr2k", [GB_RK_SSYR2K] = "ssyr2k",
    [GB_RK_DSYMM] = "dsymm",   [GB_RK_SSYMM] = "ssymm",   [GB_RK_DTRSM] = "dtrsm",
    [GB_RK_STRSM] = "strsm",   [GB_RK_DTRMM] = "dtrmm",   [GB_RK_STRMM] = "strmm",
    [GB_RK_DGEMV] = "dgemv",   [GB_RK_DGER] = "dger",     [GB_RK_DTRSV] = "dtrsv",
    [GB_RK_DSYMV] = "dsymv",   [GB_RK_DSYR] = "dsyr",     [GB_RK_STREAM] = "daxpy",
    [GB_RK_REDUCE] = "ddot",
};

struct xover {
    double flops;     /* GPU only from this many flops ... */
    double intensity; /* ... and this many flops per operand byte */
};

/* Built-in table, measured with gpublas_tune_crossover(path, 2048) on an
 * MI355X host (16 host threads, scipy's OpenBLAS): the square size from
 * which the GPU, copying its operands over the host link on every call
 * (staged through pinned memory), beat the host BLAS at every larger size
 * measured.  Level-1/2 calls never
 * won (copying the operands costs more than the host's one pass over them).
 * Other machines measure their own table and load it with xover=<file>. */
#define SQ3(a, n, es) {(a) * (double)(n) * (n) * (n), (a) * (double)(n) / (es)}
#define INF_ HUGE_VAL
static struct xover g_xover[GB_RK_COUNT] = {
    [GB_RK_DGEMM] = SQ3(2, 1024, 32),  /* 2n^3 flops over 4 n^2 doubles */
    [GB_RK_SGEMM] = SQ3(2, 1024, 16),
    [GB_RK_DSYRK] = SQ3(1, 512, 24),   /* n^3 flops over 3 n^2 doubles */
    [GB_RK_SSYRK] = SQ3(1, 768, 12),
    [GB_RK_DSYR2K] = SQ3(2, 384, 32),
    [GB_RK_SSYR2K] = SQ3(2, 384, 16),
    [GB_RK_DSYMM] = SQ3(2, 1024, 32),
    [GB_RK_SSYMM] = SQ3(2, 1024, 16),
    [GB_RK_DTRSM] = SQ3(1, 1280, 24),
    [GB_RK_STRSM] = SQ3(1, 1280, 12),
    [GB_RK_DTRMM] = SQ3(1, 1280, 24),
    [GB_RK_STRMM] = SQ3(1, 1280, 12),
    [GB_RK_DGEMV] = {INF_, INF_},
    [GB_RK_DGER] = {INF_, INF_},
    [GB_RK_DTRSV] = {INF_, INF_},
    [GB_RK_DSYMV] = {INF_, INF_},
    [GB_RK_DSYR] = {INF_, INF_},
    [GB_RK_STREAM] = {INF_, INF_},
    [GB_RK_REDUCE] = {INF_, INF_},
};

static int kind_of(const char *name)
{
    for (int i = 0; i < GB_RK_COUNT; ++i)
        if (!strcmp(name, k_names[i])) return i;
    return -1;
}

static int gpu_wins(int kind, double flops, double bytes)
{
    const struct xover *x = &g_xover[kind];
    /* the 1e-9 slack keeps a problem of exactly the measured crossover shape
     * on the GPU despite rounding in flops / bytes */
    return flops >= x->flops && flops >= x->intensity * bytes * (1.0 - 1e-9);
}

int gb_host_route(int kind, int below_fixed, double flops, double bytes, const void *p,
                  const void *q, const void *r)
{
    if (gb_forked_child) return 1; /* no GPU runtime in a fork child */
    if (gb_route_gpu) return 0;
    const int cut = gb_opt.cutoff;
    if (cut == 0) return 0;
    if (cut > 0 && !below_fixed) return 0;
    if (gb_ptr_tracked(p) || gb_ptr_tracked(q) || gb_ptr_tracked(r)) return 0;
    if (cut > 0) return 1;
    return !gpu_wins(kind, flops, bytes);
}

int gpublas_set_crossover(const char *routine, double min_flops, double min_intensity)
{
    const int k = routine ? kind_of(routine) : -1;
    if (k < 0 || !(min_flops >= 0) || !(min_intensity >= 0)) return -1;
    g_xover[k].flops = min_flops;
    g_xover[k].intensity = min_intensity;
    return 0;
}

int gpublas_get_crossover(const char *routine, double *min_flops, double *min_intensity)
{
    const int k = routine ? kind_of(routine) : -1;
    if (k < 0) return -1;
    if (min_flops) *min_flops = g_xover[k].flops;
    if (min_intensity) *min_intensity = g_xover[k].intensity;
    return 0;
}

int gpublas_crossover_gpu(const char *routine, double flops, double bytes)
{
    const int k = routine ? kind_of(routine) : -1;
    if (k < 0) return -1;
    return gpu_wins(k, flops, bytes);
}

/* Text table: '#' starts a comment; every other non-blank line is
 * "<routine> <min_flops> <min_intensity>" ("inf" allowed).  Unknown routines
 * and malformed lines make the whole file fail and leave the table as it was. */
int gb_xover_load(const char *path)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    struct xover tmp[GB_RK_COUNT];
    memcpy(tmp, g_xover, sizeof tmp);
    char line[512];
    int bad = 0, nrows = 0;
    while (fgets(line, sizeof line, f)) {
        char *h = strchr(line, '#');
        if (h) *h = 0;
        char name[32];
        char fs[64], is[64];
        int got = sscanf(line, "%31s %63s %63s", name, fs, is);
        if (got <= 0) continue;
        int k = kind_of(name);
        char *e1, *e2;
        double fl = got == 3 ? strtod(fs, &e1) : -1, in = got == 3 ? strtod(is, &e2) : -1;
        if (k < 0 || got != 3 || *e1 || *e2 || !(fl >= 0) || !(in >= 0)) {
            bad = 1;
            break;
        }
        tmp[k].flops = fl;
        tmp[k].intensity = in;
        nrows++;
    }
    fclose(f);
    if (bad || nrows == 0) return -1;
    memcpy(g_xover, tmp, sizeof tmp);
    return 0;
}

/* ------------------------------------------------------------------ */
/* measurement                                                         */
/* ------------------------------------------------------------------ */
/* the library's own entry points, bound inside the library so the device
 * side of a measurement never resolves to another BLAS of the process */
extern __typeof__(dgemm_) gb_self_dgemm_;
extern __typeof__(sgemm_) gb_self_sgemm_;
extern __typeof__(dsyrk_) gb_self_dsyrk_;
extern __typeof__(ssyrk_) gb_self_ssyrk_;
extern __typeof__(dsyr2k_) gb_self_dsyr2k_;
extern __typeof__(ssyr2k_) gb_self_ssyr2k_;
extern __typeof__(dsymm_) gb_self_dsymm_;
extern __typeof__(ssymm_) gb_self_ssymm_;
extern __typeof__(dtrsm_) gb_self_dtrsm_;
extern __typeof__(strsm_) gb_self_strsm_;
extern __typeof__(dtrmm_) gb_self_dtrmm_;
extern __typeof__(strmm_) gb_self_strmm_;
extern __typeof__(dgemv_) gb_self_dgemv_;
extern __typeof__(dger_) gb_self_dger_;
extern __typeof__(dtrsv_) gb_self_dtrsv_;
extern __typeof__(dsymv_) gb_self_dsymv_;
extern __typeof__(dsyr_) gb_self_dsyr_;
extern __typeof__(daxpy_) gb_self_daxpy_;
extern __typeof__(ddot_) gb_self_ddot_;

static const char *const k_f77[GB_RK_COUNT] = {
    [GB_RK_DGEMM] = "dgemm_",   [GB_RK_SGEMM] = "sgemm_",   [GB_RK_DSYRK] = "dsyrk_",
    [GB_RK_SSYRK] = "ssyrk_",   [GB_RK_DSYR2K] = "dsyr2k_", [GB_RK_SSYR2K] = "ssyr2k_",
    [GB_RK_DSYMM] = "dsymm_",   [GB_RK_SSYMM] = "ssymm_",   [GB_RK_DTRSM] = "dtrsm_",
    [GB_RK_STRSM] = "strsm_",   [GB_RK_DTRMM] = "dtrmm_",   [GB_RK_STRMM] = "strmm_",
    [GB_RK_DGEMV] = "dgemv_",   [GB_RK_DGER] = "dger_",     [GB_RK_DTRSV] = "dtrsv_",
    [GB_RK_DSYMV] = "dsymv_",   [GB_RK_DSYR] = "dsyr_",     [GB_RK_STREAM] = "daxpy_",
    [GB_RK_REDUCE] = "ddot_",
};

static const void *self_fn(int k)
{
    switch (k) {
    case GB_RK_DGEMM: return (const void *)gb_self_dgemm_;
    case GB_RK_SGEMM: return (const void *)gb_self_sgemm_;
    case GB_RK_DSYRK: return (const void *)gb_self_dsyrk_;
    case GB_RK_SSYRK: return (const void *)gb_self_ssyrk_;
    case GB_RK_DSYR2K: return (const void *)gb_self_dsyr2k_;
    case GB_RK_SSYR2K: return (const void *)gb_self_ssyr2k_;
    case GB_RK_DSYMM: return (const void *)gb_self_dsymm_;
    case GB_RK_SSYMM: return (const void *)gb_self_ssymm_;
    case GB_RK_DTRSM: return (const void *)gb_self_dtrsm_;
    case GB_RK_STRSM: return (const void *)gb_self_strsm_;
    case GB_RK_DTRMM: return (const void *)gb_self_dtrmm_;
    case GB_RK_STRMM: return (const void *)gb_self_strmm_;
    case GB_RK_DGEMV: return (const void *)gb_self_dgemv_;
    case GB_RK_DGER: return (const void *)gb_self_dger_;
    case GB_RK_DTRSV: return (const void *)gb_self_dtrsv_;
    case GB_RK_DSYMV: return (const void *)gb_self_dsymv_;
    case GB_RK_DSYR: return (const void *)gb_self_dsyr_;
    case GB_RK_STREAM: return (const void *)gb_self_daxpy_;
    default: return (const void *)gb_self_ddot_;
    }
}

static int is_single(int k)
{
    return k == GB_RK_SGEMM || k == GB_RK_SSYRK || k == GB_RK_SSYR2K || k == GB_RK_SSYMM ||
           k == GB_RK_STRSM || k == GB_RK_STRMM;
}

static int level_of(int k) { return k <= GB_RK_STRMM ? 3 : k <= GB_RK_DSYR ? 2 : 1; }

/* work model of a square problem of size n: the same formulas the entry
 * points apply to general shapes */
void gb_route_model(int k, double n, double *flops, double *bytes)
{
    const double es = is_single(k) ? 4 : 8;
    double f, b;
    switch (k) {
    case GB_RK_DGEMM: case GB_RK_SGEMM: /* 2mnk; A, B, C in and out */
    case GB_RK_DSYR2K: case GB_RK_SSYR2K: /* 2n^2k; A, B, C in and out */
    case GB_RK_DSYMM: case GB_RK_SSYMM: /* 2mn*na; A, B, C in and out */
        f = 2 * n * n * n, b = es * 4 * n * n; break;
    case GB_RK_DSYRK: case GB_RK_SSYRK: /* n^2k; A, C in and out */
    case GB_RK_DTRSM: case GB_RK_STRSM: /* mn*na; A, B in and out */
    case GB_RK_DTRMM: case GB_RK_STRMM:
        f = n * n * n, b = es * 3 * n * n; break;
    case GB_RK_DGEMV: f = 2 * n * n, b = es * (n * n + 2 * n); break; /* A, x, y */
    case GB_RK_DGER: f = 2 * n * n, b = es * (2 * n * n + 2 * n); break; /* A in and out */
    case GB_RK_DTRSV: f = n * n, b = es * (n * n + 2 * n); break;
    case GB_RK_DSYMV: f = 2 * n * n, b = es * (n * n + 3 * n); break; /* A, x, y in and out */
    case GB_RK_DSYR: f = n * n, b = es * (2 * n * n + n); break;      /* A in and out, x */
    case GB_RK_STREAM: f = 2 * n, b = es * 3 * n; break; /* x, y in and out */
    default: f = 2 * n, b = es * 2 * n; break;           /* x, y */
    }
    *flops = f;
    *bytes = b;
}

static double now_us(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec * 1e6 + t.tv_nsec * 1e-3;
}

struct bufs {
    void *a, *b, *c, *b0, *c0; /* operands and pristine copies of the outputs */
    size_t mat_bytes;
};

/* one call of routine k on size n through fn (host BLAS or the library) */
static void call_once(int k, int n, const void *fn, struct bufs *u)
{
    const int one = 1;
    const double d1 = 1.0, dh = 0.5;
    const float s1 = 1.0f, sh = 0.5f;
    const char N = 'N', L = 'L';
    typedef void (*gemm_d)(const char *, const char *, const int *, const int *, const int *,
                           const double *, const double *, const int *, const double *, const int *,
                           const double *, double *, const int *);
    typedef void (*gemm_s)(const char *, const char *, const int *, const int *, const int *,
                           const float *, const float *, const int *, const float *, const int *,
                           const float *, float *, const int *);
    typedef void (*syrk_d)(const char *, const char *, const int *, const int *, const double *,
                           const double *, const int *, const double *, double *, const int *);
    typedef void (*syrk_s)(const char *, const char *, const int *, const int *, const float *,
                           const float *, const int *, const float *, float *, const int *);
    typedef void (*syr2k_d)(const char *, const char *, const int *, const int *, const double *,
                            const double *, const int *, const double *, const int *,
                            const double *, double *, const int *);
    typedef void (*syr2k_s)(const char *, const char *, const int *, const int *, const float *,
                            const float *, const int *, const float *, const int *, const float *,
                            float *, const int *);
    typedef void (*tr_d)(const char *, const char *, const char *, const char *, const int *,
                         const int *, const double *, const double *, const int *, double *,
                         const int *);
    typedef void (*tr_s)(const char *, const char *, const char *, const char *, const int *,
                         const int *, const float *, const float *, const int *, float *,
                         const int *);
    typedef void (*gemv_t)(const char *, const int *, const int *, const double *, const double *,
                           const int *, const double *, const int *, const double *, double *,
                           const int *);
    typedef void (*ger_t)(const int *, const int *, const double *, const double *, const int *,
                          const double *, const int *, double *, const int *);
    typedef void (*trsv_t)(const char *, const char *, const char *, const int *, const double *,
                           const int *, double *, const int *);
    typedef void (*axpy_t)(const int *, const double *, const double *, const int *, double *,
                           const int *);
    typedef double (*dot_t)(const int *, const double *, const int *, const double *,
                            const int *);
    typedef void (*symv_t)(const char *, const int *, const double *, const double *,
                           const int *, const double *, const int *, const double *, double *,
                           const int *);
    typedef void (*syr_t)(const char *, const int *, const double *, const double *, const int *,
                          double *, const int *);
    switch (k) {
    case GB_RK_DGEMM:
        ((gemm_d)fn)(&N, &N, &n, &n, &n, &d1, u->a, &n, u->b, &n, &d1, u->c, &n); break;
    case GB_RK_SGEMM:
        ((gemm_s)fn)(&N, &N, &n, &n, &n, &s1, u->a, &n, u->b, &n, &s1, u->c, &n); break;
    case GB_RK_DSYRK: ((syrk_d)fn)(&L, &N, &n, &n, &d1, u->a, &n, &d1, u->c, &n); break;
    case GB_RK_SSYRK: ((syrk_s)fn)(&L, &N, &n, &n, &s1, u->a, &n, &s1, u->c, &n); break;
    case GB_RK_DSYR2K:
        ((syr2k_d)fn)(&L, &N, &n, &n, &d1, u->a, &n, u->b, &n, &d1, u->c, &n); break;
    case GB_RK_SSYR2K:
        ((syr2k_s)fn)(&L, &N, &n, &n, &s1, u->a, &n, u->b, &n, &s1, u->c, &n); break;
    case GB_RK_DSYMM:
        ((syr2k_d)fn)(&L, &L, &n, &n, &d1, u->a, &n, u->b, &n, &d1, u->c, &n); break;
    case GB_RK_SSYMM:
        ((syr2k_s)fn)(&L, &L, &n, &n, &s1, u->a, &n, u->b, &n, &s1, u->c, &n); break;
    case GB_RK_DTRSM:
    case GB_RK_DTRMM:
        ((tr_d)fn)(&L, &L, &N, &N, &n, &n, &dh, u->a, &n, u->b, &n); break;
    case GB_RK_STRSM:
    case GB_RK_STRMM:
        ((tr_s)fn)(&L, &L, &N, &N, &n, &n, &sh, u->a, &n, u->b, &n); break;
    case GB_RK_DGEMV:
        ((gemv_t)fn)(&N, &n, &n, &d1, u->a, &n, u->b, &one, &d1, u->c, &one); break;
    case GB_RK_DGER: ((ger_t)fn)(&n, &n, &dh, u->b, &one, u->c, &one, u->a, &n); break;
    case GB_RK_DTRSV: ((trsv_t)fn)(&L, &N, &N, &n, u->a, &n, u->b, &one); break;
    case GB_RK_DSYMV: ((symv_t)fn)(&L, &n, &d1, u->a, &n, u->b, &one, &d1, u->c, &one); break;
    case GB_RK_DSYR: ((syr_t)fn)(&L, &n, &dh, u->b, &one, u->a, &n); break;
    case GB_RK_STREAM: ((axpy_t)fn)(&n, &dh, u->a, &one, u->b, &one); break;
    default: {
        volatile double r = ((dot_t)fn)(&n, u->a, &one, u->b, &one);
        (void)r;
    }
    }
}

/* restore what a call overwrites, so every timed call sees the same data */
static void reset(int k, int n, struct bufs *u)
{
    const size_t es = is_single(k) ? 4 : 8;
    const int lvl = level_of(k);
    const size_t len = lvl == 1 ? (size_t)n : (size_t)n * n;
    if (k == GB_RK_DGER) memcpy(u->a, u->b0, len * es);
    else if (k == GB_RK_DTRSM || k == GB_RK_STRSM || k == GB_RK_DTRMM || k == GB_RK_STRMM)
        memcpy(u->b, u->b0, len * es);
    else if (k == GB_RK_DTRSV) memcpy(u->b, u->b0, (size_t)n * es);
    else if (k == GB_RK_STREAM) memcpy(u->b, u->b0, len * es);
    else if (k == GB_RK_DGEMV || k == GB_RK_DSYMV) memcpy(u->c, u->c0, (size_t)n * es);
    else if (k == GB_RK_DSYR) return; /* grows by alpha*x*x' per call: bounded */
    else if (k != GB_RK_REDUCE) memcpy(u->c, u->c0, len * es);
}

/* best time of several calls (each after a reset), first call untimed */
static double best_us(int k, int n, const void *fn, struct bufs *u, int gpu)
{
    double best = HUGE_VAL, spent = 0;
    for (int rep = 0; rep < 40; ++rep) {
        reset(k, n, u);
        gb_route_gpu = gpu;
        const double t0 = now_us();
        call_once(k, n, fn, u);
        const double dt = now_us() - t0;
        gb_route_gpu = 0;
        if (rep == 0) continue;
        if (dt < best) best = dt;
        spent += dt;
        if (rep >= 3 && spent > 20000) break;
    }
    return best;
}

static void fill(void *p, size_t count, int single, unsigned seed, double diag_boost, int n)
{
    /* values in [-1/2, 1/2); for triangular solves a dominant diagonal keeps
     * the solution bounded */
    unsigned s = seed * 2654435761u + 12345u;
    for (size_t i = 0; i < count; ++i) {
        s = s * 1664525u + 1013904223u;
        double v = ((s >> 8) & 0xffff) / 65536.0 - 0.5;
        if (diag_boost != 0 && n > 0 && i % ((size_t)n + 1) == 0) v = diag_boost;
        if (single) ((float *)p)[i] = (float)v;
        else ((double *)p)[i] = v;
    }
}

static const int k_ladder3[] = {8,   16,  24,  32,  48,  64,  96,   128,  160,  192, 256,
                                320, 384, 512, 640, 768, 1024, 1280, 1536, 2048, 3072, 4096};
static const int k_ladder2[] = {64, 128, 256, 512, 1024, 2048, 4096, 8192};
static const int k_ladder1[] = {1 << 10, 1 << 12, 1 << 14, 1 << 16, 1 << 18,
                                1 << 20, 1 << 22, 1 << 24, 1 << 26};

int gpublas_tune_crossover(const char *path, int max_n)
{
    if (max_n < 16) return -1;
    gb_enter(); /* buffers allocated here are never tracked */
    gb_require_device();
    const int max3 = max_n, max2 = 2 * max_n;
    const size_t max1 = (size_t)max_n * 4096;
    const size_t mat = (size_t)max2 * max2 > max1 ? (size_t)max2 * max2 : max1;
    struct bufs u = {0};
    u.mat_bytes = mat * 8;
    void **all[5] = {&u.a, &u.b, &u.c, &u.b0, &u.c0};
    for (int i = 0; i < 5; ++i) {
        *all[i] = malloc(u.mat_bytes);
        if (!*all[i]) {
            for (int j = 0; j < i; ++j) free(*all[j]);
            gb_leave();
            return -2;
        }
    }
    FILE *f = path ? fopen(path, "w") : NULL;
    if (path && !f) {
        for (int j = 0; j < 5; ++j) free(*all[j]);
        gb_leave();
        return -3;
    }
    if (f) {
        Dl_info di;
        const char *host = dladdr(gb_host_fn("dgemm_"), &di) && di.dli_fname ? di.dli_fname : "?";
        const char *thr = getenv("OMP_NUM_THREADS");
        fprintf(f, "# libgpublas host/GPU crossover table (gpublas_tune_crossover, max_n %d)\n",
                max_n);
        fprintf(f, "# host BLAS %s, OMP_NUM_THREADS=%s\n", host, thr ? thr : "unset");
        fprintf(f, "# square problems of size n, operands not resident (copied over the host\n"
                   "# link on every device call), best of several calls per side\n");
        fprintf(f, "# routine n flops bytes host_us gpu_us\n");
    }
    struct xover res[GB_RK_COUNT];
    for (int k = 0; k < GB_RK_COUNT; ++k) {
        const int lvl = level_of(k), single = is_single(k);
        const int *lad = lvl == 3 ? k_ladder3 : lvl == 2 ? k_ladder2 : k_ladder1;
        const int nl = lvl == 3   ? (int)(sizeof k_ladder3 / sizeof *k_ladder3)
                       : lvl == 2 ? (int)(sizeof k_ladder2 / sizeof *k_ladder2)
                                  : (int)(sizeof k_ladder1 / sizeof *k_ladder1);
        const size_t cap = lvl == 3 ? (size_t)max3 : lvl == 2 ? (size_t)max2 : max1;
        const void *host = gb_host_fn(k_f77[k]);
        const void *dev = self_fn(k);
        int ns[32], cnt = 0;
        for (int i = 0; i < nl && (size_t)lad[i] <= cap && cnt < 32; ++i) ns[cnt++] = lad[i];
        double th[32], tg[32];
        /* Each side is timed in a phase of its own, the device first after a
         * pause: a host BLAS's threads keep spinning for a while after a call
         * and would otherwise take the cores the device path's copies use. */
        const struct timespec pause = {0, 200 * 1000 * 1000};
        nanosleep(&pause, NULL);
        for (int side = 1; side >= 0; --side)
            for (int i = 0; i < cnt; ++i) {
                const int n = ns[i];
                const size_t len = lvl == 1 ? (size_t)n : (size_t)n * n;
                const int tri = k == GB_RK_DTRSM || k == GB_RK_STRSM || k == GB_RK_DTRMM ||
                                k == GB_RK_STRMM || k == GB_RK_DTRSV;
                /* TRSM divides by a diagonal of 1 + n/2: the solve stays near
                 * the right-hand side's scale; TRMM keeps the same matrix */
                fill(u.a, len, single, 1, tri ? 1.0 + n / 2.0 : 0.0, lvl == 1 ? 0 : n);
                fill(u.b0, len, single, 2, 0, 0);
                fill(u.c0, len, single, 3, 0, 0);
                memcpy(u.b, u.b0, len * (single ? 4 : 8));
                if (side) tg[i] = best_us(k, n, dev, &u, 1);
                else th[i] = best_us(k, n, host, &u, 0);
            }
        int run_start = -1;
        for (int i = 0; i < cnt; ++i) {
            double fl, by;
            gb_route_model(k, ns[i], &fl, &by);
            if (f) {
                fprintf(f, "# %s %d %.6g %.6g %.2f %.2f\n", k_names[k], ns[i], fl, by, th[i],
                        tg[i]);
                fflush(f); /* a long measurement shows its progress in the file */
            }
            if (tg[i] < th[i]) {
                if (run_start < 0) run_start = ns[i];
            } else {
                run_start = -1;
            }
        }
        res[k].flops = res[k].intensity = HUGE_VAL;
        if (run_start > 0) {
            double fl, by;
            gb_route_model(k, run_start, &fl, &by);
            res[k].flops = fl;
            res[k].intensity = fl / by;
        }
    }
    if (f) {
        fprintf(f, "# routine min_flops min_intensity (GPU when flops >= min_flops and\n"
                   "# flops / operand bytes >= min_intensity; inf = the host at every size)\n");
        for (int k = 0; k < GB_RK_COUNT; ++k)
            fprintf(f, "%s %.9g %.9g\n", k_names[k], res[k].flops, res[k].intensity);
        fclose(f);
    }
    memcpy(g_xover, res, sizeof res);
    for (int j = 0; j < 5; ++j) free(*all[j]);
    gb_leave();
    return 0;
}
