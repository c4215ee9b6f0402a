/*
 * core.c -- library lifecycle, options, statistics, xerbla/lsame and the
 * host-BLAS fall-through resolver.
 *
 * Reference counterparts: blas2cuda.c:46-124 (options), :178-277
 * (init/fini, statistics.csv), runtime-blas.c:34-69 (xerbla, lsame,
 * RTLD_NEXT lookup), common.h:15-28 (malloc-free logging).
 */
#include "internal.h"
#include <dlfcn.h>
#include <errno.h>
#include <fcntl.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#include <sys/syscall.h>

GB_TLS int gb_guard = 0;
gpublas_stats_t gb_stats;
struct gb_options gb_opt = {
    .heuristic = H_FALSE,
    .size_threshold = 1u << 20,
    .min_track = 4096,
    .cutoff = 512,
    .serve_budget = (size_t)64 << 30,
    .serve_slots = 8,
    .serve_graph = 0,
    .serve_batch = 256,
};

static int g_initialized = 0;

void gb_log(int fd, const char *fmt, ...)
{
    char buf[2048];
    va_list ap;
    va_start(ap, fmt);
    int n = vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (n < 0) return;
    if ((size_t)n >= sizeof buf) n = sizeof buf - 1;
    ssize_t off = 0;
    while (off < n) {
        ssize_t w = syscall(SYS_write, fd, buf + off, (size_t)(n - off));
        if (w <= 0) break;
        off += w;
    }
}

void gb_fatal(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    gb_log(2, "libgpublas: fatal: %s\n", buf);
    abort();
}

void gb_check(int err, const char *what)
{
    if (err) gb_fatal("%s failed (HIP error %d)", what, err);
}

/* ------------------------------------------------------------------ */
/* options                                                             */
/* ------------------------------------------------------------------ */
static void print_help(void)
{
    gb_log(2,
           "libgpublas options (GPUBLAS_OPTIONS, ';'-separated):\n"
           "  help                 print this help\n"
           "  heuristic=<h>        which allocations become device-resident objects:\n"
           "                       false (default) | true | random | size:<bytes> |\n"
           "                       objtrack:<file> (alias oracle:<file>)\n"
           "  min_track=<bytes>    never track allocations smaller than this (4096)\n"
           "  cutoff=<n>           GEMM leading dimensions (Level-1/2 lengths) below n go\n"
           "                       to the host BLAS when no operand is resident (512;\n"
           "                       0 = always the GPU)\n"
           "  cutoff=auto          host or GPU by the measured crossover table\n"
           "  xover=<file>         crossover table written by gpublas_tune_crossover\n"
           "  profile              time every kernel with HIP events\n"
           "  split=<s>            DGEMM/DSYRK as split-precision slice products\n"
           "                       (s = 1..8 slices per operand; 0 = plain FP64, default)\n"
           "  slice=<bf16|fp16|int8> slice format of split (bf16; fp16 takes s <= 6; int8:\n"
           "                       7-bit slices on integer MFMA, s = 7 for FP64-level accuracy)\n"
           "  serve[=<bytes>]      run DGEMM/SGEMM/DGEMV/left DTRSM calls whose operands are\n"
           "                       all tracked through the batching serving loop, with an\n"
           "                       HBM admission budget (64G); serve=0 turns it off\n"
           "  serve_slots=<n>      batches in flight (8)\n"
           "  serve_graph=<0|1>    replay captured hipGraphs (1) or launch eagerly (0, default)\n"
           "  resident[=0|1]       DDOT/DNRM2/IDAMAX/IDAMIN/DASUM (n <= 65536) on tracked operands\n"
           "                       through a kernel resident on the device (no launch per call)\n"
           "  serve_batch=<n>      most calls per batch (256; 1 = one launch per call)\n"
           "  tune=<file>          tile-shape table of the batch GEMM kernels: loaded\n"
           "                       from <file>, measured and written there if absent\n"
           "  stats[=<file>]       write hits/misses at exit (statistics.csv)\n"
           "  trace=<file>         log allocations (T/U) and BLAS operand uses (C)\n"
           "  objtrack_out=<file>  at exit, write an object-tracking file naming the\n"
           "                       call sites and ordinals of allocations that were BLAS\n"
           "                       operands (capture with heuristic=true)\n"
           "  verbose              log decisions to stderr\n");
}

static size_t parse_size(const char *s)
{
    char *end = NULL;
    double v = strtod(s, &end);
    if (end && (*end == 'k' || *end == 'K')) v *= 1024.0;
    else if (end && (*end == 'm' || *end == 'M')) v *= 1048576.0;
    else if (end && (*end == 'g' || *end == 'G')) v *= 1073741824.0;
    return (size_t)v;
}

int gb_parse_options(const char *s)
{
    if (!s) return 0;
    char buf[4096];
    snprintf(buf, sizeof buf, "%s", s);
    int bad = 0;
    char *save = NULL;
    for (char *opt = strtok_r(buf, ";", &save); opt; opt = strtok_r(NULL, ";", &save)) {
        while (*opt == ' ') opt++;
        if (!*opt) continue;
        if (!strcmp(opt, "help")) print_help();
        else if (!strcmp(opt, "profile")) gb_opt.profile = 1;
        else if (!strncmp(opt, "split=", 6)) {
            gb_opt.split = atoi(opt + 6);
            if (gb_opt.split < 0 || gb_opt.split > 8) {
                gb_log(2, "libgpublas: split=%d out of range (0 = off, 1..8 slices)\n", gb_opt.split);
                gb_opt.split = 0;
                bad = 1;
            }
        }
        else if (!strcmp(opt, "slice=bf16")) gb_opt.split_fmt = GPUBLAS_SLICE_BF16;
        else if (!strcmp(opt, "slice=fp16")) gb_opt.split_fmt = GPUBLAS_SLICE_FP16;
        else if (!strcmp(opt, "slice=int8")) gb_opt.split_fmt = GPUBLAS_SLICE_INT8;
        else if (!strcmp(opt, "serve")) gb_opt.serve = 1;
        else if (!strcmp(opt, "resident") || !strcmp(opt, "resident=1")) gb_opt.resident = 1;
        else if (!strcmp(opt, "resident=0")) gb_opt.resident = 0;
        else if (!strncmp(opt, "serve=", 6)) {
            const size_t b = parse_size(opt + 6);
            gb_opt.serve = b != 0;
            if (b) gb_opt.serve_budget = b;
        }
        else if (!strncmp(opt, "serve_slots=", 12)) {
            const int v = atoi(opt + 12);
            if (v >= 1 && v <= 16) gb_opt.serve_slots = v;
            else bad = 1;
            gb_serve_stop();
        }
        else if (!strncmp(opt, "serve_graph=", 12)) {
            gb_opt.serve_graph = atoi(opt + 12) != 0;
            gb_serve_stop();
        }
        else if (!strncmp(opt, "serve_batch=", 12)) {
            const int v = atoi(opt + 12);
            if (v >= 1) gb_opt.serve_batch = v;
            else bad = 1;
        }
        else if (!strncmp(opt, "tune=", 5)) {
            snprintf(gb_opt.tune_file, sizeof gb_opt.tune_file, "%s", opt + 5);
            gb_serve_stop();
        }
        else if (!strcmp(opt, "verbose")) gb_opt.verbose = 1;
        else if (!strcmp(opt, "stats")) snprintf(gb_opt.stats_file, sizeof gb_opt.stats_file, "statistics.csv");
        else if (!strncmp(opt, "stats=", 6)) snprintf(gb_opt.stats_file, sizeof gb_opt.stats_file, "%s", opt + 6);
        else if (!strncmp(opt, "trace=", 6)) snprintf(gb_opt.trace_file, sizeof gb_opt.trace_file, "%s", opt + 6);
        else if (!strncmp(opt, "objtrack_out=", 13))
            snprintf(gb_opt.objtrack_out, sizeof gb_opt.objtrack_out, "%s", opt + 13);
        else if (!strcmp(opt, "cutoff=auto")) gb_opt.cutoff = GB_CUTOFF_AUTO;
        else if (!strncmp(opt, "cutoff=", 7)) {
            char *end;
            long v = strtol(opt + 7, &end, 10);
            if (*end || v < 0 || v > 1 << 30) bad = 1;
            else gb_opt.cutoff = (int)v;
        }
        else if (!strncmp(opt, "xover=", 6)) {
            if (gb_xover_load(opt + 6) != 0) {
                gb_log(2, "libgpublas: cannot load crossover table '%s'\n", opt + 6);
                bad = 1;
            }
        }
        else if (!strncmp(opt, "min_track=", 10)) gb_opt.min_track = parse_size(opt + 10);
        else if (!strncmp(opt, "heuristic=", 10)) {
            const char *h = opt + 10;
            if (!strcmp(h, "true")) gb_opt.heuristic = H_TRUE;
            else if (!strcmp(h, "false")) gb_opt.heuristic = H_FALSE;
            else if (!strcmp(h, "random")) gb_opt.heuristic = H_RANDOM;
            else if (!strncmp(h, "size:", 5)) {
                gb_opt.heuristic = H_SIZE;
                gb_opt.size_threshold = parse_size(h + 5);
            } else if (!strncmp(h, "objtrack:", 9) || !strncmp(h, "oracle:", 7)) {
                const char *file = strchr(h, ':') + 1;
                if (gb_objtrack_load(file) != 0) {
                    gb_log(2, "libgpublas: cannot load object-tracking file '%s'\n", file);
                    bad = 1;
                } else {
                    gb_opt.heuristic = H_OBJTRACK;
                }
            } else {
                gb_log(2, "libgpublas: unsupported heuristic '%s'\n", h);
                bad = 1;
            }
        } else {
            gb_log(2, "libgpublas: unknown option '%s' (GPUBLAS_OPTIONS=help)\n", opt);
            bad = 1;
        }
    }
    if (gb_opt.split_fmt == GPUBLAS_SLICE_FP16 && gb_opt.split > 6) {
        gb_log(2, "libgpublas: slice=fp16 takes at most 6 slices; using split=6\n");
        gb_opt.split = 6;
        bad = 1;
    }
    return bad;
}

/* ------------------------------------------------------------------ */
/* lifecycle                                                           */
/* ------------------------------------------------------------------ */
__attribute__((constructor)) void gb_init(void)
{
    if (g_initialized) return;
    g_initialized = 1;
    gb_enter();
    const char *o = getenv("GPUBLAS_OPTIONS");
    if (o) gb_parse_options(o);
    gb_install_fault_handler();
    if (gb_opt.profile) gpublas_dev_prof_enable(1);
    gb_tracking = 1;
    gb_leave();
}

__attribute__((destructor)) static void gb_fini(void)
{
    gb_enter();
    gb_tracking = 0;
    if (gb_forked_child) { /* the serving thread, the files and the GPU are the parent's */
        gb_leave();
        return;
    }
    gb_serve_stop();
    if (gb_opt.objtrack_out[0]) gb_objtrack_write();
    if (gb_opt.stats_file[0]) {
        int fd = open(gb_opt.stats_file, O_WRONLY | O_CREAT | O_TRUNC, 0644);
        if (fd >= 0) {
            gb_log(fd, "Hits, Misses, DeviceCalls, Fallthrough, Faults, Uploads\n");
            gb_log(fd, "%llu, %llu, %llu, %llu, %llu, %llu\n",
                   (unsigned long long)gb_stats.hits, (unsigned long long)gb_stats.misses,
                   (unsigned long long)gb_stats.device_calls,
                   (unsigned long long)gb_stats.fallthrough, (unsigned long long)gb_stats.faults,
                   (unsigned long long)gb_stats.uploads);
            close(fd);
        }
    }
    gb_leave();
}

/* Running the shared object prints the help (reference entry.c:4-11). */
#ifdef GPUBLAS_EXEC_ENTRY
const char gb_interp[] __attribute__((section(".interp"))) = "/lib64/ld-linux-x86-64.so.2";
void gb_entry(void)
{
    print_help();
    _exit(0);
}
#endif

/* ------------------------------------------------------------------ */
/* xerbla / lsame                                                      */
/* ------------------------------------------------------------------ */
extern void xerbla_(const char *name, const int *info, size_t len) __attribute__((weak));
static GB_TLS int t_last_info = 0;
static GB_TLS char t_last_name[8];

int gb_cblas_translate(int info);

void gb_xerbla(const char *name6, int info)
{
    info = gb_cblas_translate(info);
    t_last_info = info;
    memcpy(t_last_name, name6, 6);
    t_last_name[6] = 0;
    if (xerbla_) {
        xerbla_(name6, &info, 6);
        return;
    }
    /* netlib XERBLA wording; netlib then STOPs, a library returns instead */
    char nm[7];
    memcpy(nm, name6, 6);
    nm[6] = 0;
    for (int i = 5; i >= 0 && nm[i] == ' '; --i) nm[i] = 0;
    gb_log(2, " ** On entry to %s parameter number %2d had an illegal value\n", nm, info);
}

int gpublas_last_xerbla(char name[8])
{
    if (name) memcpy(name, t_last_name, 8);
    return t_last_info;
}

int gb_lsame(char a, char b)
{
    if (a >= 'a' && a <= 'z') a -= 32;
    if (b >= 'a' && b <= 'z') b -= 32;
    return a == b;
}

int lsame_(const char *ca, const char *cb) { return gb_lsame(*ca, *cb); }

/* ------------------------------------------------------------------ */
/* host BLAS                                                           */
/* ------------------------------------------------------------------ */
static pthread_mutex_t g_host_mu = PTHREAD_MUTEX_INITIALIZER;
static void *g_host_lib = NULL;
static char g_host_prefix[32];
static int g_host_tried = 0;

static void open_host_lib(void)
{
    const char *env = getenv("GPUBLAS_HOST_BLAS");
    const char *pre = getenv("GPUBLAS_HOST_BLAS_PREFIX");
    const char *cands[] = {
        env,
        "libopenblas.so.0",
        "libblas.so.3",
        "libblas.so",
        "/usr/local/lib/python3.10/dist-packages/scipy.libs/libscipy_openblas-68440149.so",
    };
    for (size_t i = 0; i < sizeof cands / sizeof cands[0]; ++i) {
        if (!cands[i] || !*cands[i]) continue;
        void *h = dlopen(cands[i], RTLD_NOW | RTLD_LOCAL);
        if (!h) continue;
        g_host_lib = h;
        if (pre) snprintf(g_host_prefix, sizeof g_host_prefix, "%s", pre);
        else if (!dlsym(h, "dgemm_") && dlsym(h, "scipy_dgemm_"))
            snprintf(g_host_prefix, sizeof g_host_prefix, "scipy_");
        if (gb_opt.verbose) gb_log(2, "libgpublas: host BLAS %s (prefix '%s')\n", cands[i], g_host_prefix);
        return;
    }
}

void *gb_host_fn(const char *f77name)
{
    void *f = dlsym(RTLD_NEXT, f77name);
    if (f) return f;
    pthread_mutex_lock(&g_host_mu);
    if (!g_host_tried) {
        g_host_tried = 1;
        open_host_lib();
    }
    pthread_mutex_unlock(&g_host_mu);
    if (g_host_lib) {
        char nm[96];
        snprintf(nm, sizeof nm, "%s%s", g_host_prefix, f77name);
        f = dlsym(g_host_lib, nm);
    }
    if (!f) gb_fatal("no host BLAS provides '%s' (set GPUBLAS_HOST_BLAS)", f77name);
    return f;
}

/* ------------------------------------------------------------------ */
/* device                                                              */
/* ------------------------------------------------------------------ */
void gb_require_device(void)
{
    if (gb_forked_child)
        gb_fatal("BLAS on the GPU is not available in a child process after fork(); the "
                 "child's calls run on the host BLAS");
    gb_check_fault_handler();
    int st = gpublas_dev_init();
    if (st != 0)
        gb_fatal("no usable gfx950 GPU (device init status %d); libgpublas runs its BLAS "
                 "kernels on an MI355X only",
                 st);
}

int gpublas_device_ready(void) { return gpublas_dev_init() == 0; }

/* ------------------------------------------------------------------ */
/* extension API                                                       */
/* ------------------------------------------------------------------ */
void gpublas_get_stats(gpublas_stats_t *out)
{
    if (!out) return;
    *out = gb_stats;
    out->tracked_objects = (uint64_t)gb_tracked_count();
    out->allocs_seen = gb_nallocs;
    out->resident_bytes = __atomic_load_n(&gb_resident_bytes, __ATOMIC_RELAXED);
}

void gpublas_sync(void)
{
    gb_enter();
    if (gpublas_dev_init() == 0) gb_check(gpublas_dev_sync(NULL), "thread synchronisation");
    gb_leave();
}

void gpublas_reset_stats(void)
{
    uint64_t to = gb_stats.tracked_objects, tb = gb_stats.tracked_bytes;
    memset(&gb_stats, 0, sizeof gb_stats);
    gb_stats.tracked_objects = to;
    gb_stats.tracked_bytes = tb;
}

int gpublas_set_options(const char *opts)
{
    gb_enter();
    int r = gb_parse_options(opts);
    if (gb_opt.profile) gpublas_dev_prof_enable(1);
    gb_leave();
    return r;
}

void gpublas_set_cutoff(int cutoff) { gb_opt.cutoff = cutoff < 0 ? GB_CUTOFF_AUTO : cutoff; }
int gpublas_get_cutoff(void) { return gb_opt.cutoff; }

int gpublas_kernel_time(const char *kernel, double *total_ms)
{
    gb_enter();
    int n = gpublas_dev_prof_query(kernel, total_ms, NULL, NULL);
    gb_leave();
    return n;
}
