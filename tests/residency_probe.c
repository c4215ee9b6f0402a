/*
 * residency_probe.c -- an ordinary program run under
 * LD_PRELOAD=libgpublas_preload.so that checks tracked memory behaves like
 * ordinary memory towards system calls, stdio, signal handlers and fork.
 *
 *   residency_probe cpu   objects put in SHARED / DEVICE state with the
 *                         library's test hook (no GPU, no device copy)
 *   residency_probe gpu   objects made device-newer by real BLAS calls
 *                         (dscal_ on the GPU, GPUBLAS_OPTIONS=cutoff=0)
 *
 * Prints "ok <case>" per case and "FAIL <case>: ..." on a mismatch; the exit
 * status is the number of failures.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <errno.h>
#include <fcntl.h>
#include <pthread.h>
#include <setjmp.h>
#include <signal.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/uio.h>
#include <sys/wait.h>
#include <unistd.h>

enum { HOST = 0, SHARED = 1, DEVICE = 2 };

static void *(*g_malloc)(size_t);
static void (*g_free)(void *);
static int (*g_set)(void *, int);
static int (*g_get)(const void *);
static void (*g_dscal)(const int *, const double *, double *, const int *);
static double (*g_ddot)(const int *, const double *, const int *, const double *, const int *);
static int g_gpu = 0;
static int g_fail = 0;
static char g_dir[256];

#define N (1 << 17) /* doubles per object: 1 MiB */

static void check(int ok, const char *name, const char *what)
{
    if (ok) {
        printf("ok %s\n", name);
    } else {
        printf("FAIL %s: %s\n", name, what);
        g_fail++;
    }
    fflush(stdout);
}

static void *sym(const char *n)
{
    void *f = dlsym(RTLD_DEFAULT, n);
    if (!f) {
        printf("FAIL setup: no symbol %s (run under LD_PRELOAD=libgpublas_preload.so)\n", n);
        exit(99);
    }
    return f;
}

/* a tracked vector holding v[i] = scale * i, made device-newer: on the GPU
 * by dscal_ (its values are then 2 * scale * i), else by the test hook */
static double *device_newer(double scale, double *expect_factor)
{
    double *v = g_malloc(N * sizeof(double));
    for (int i = 0; i < N; ++i) v[i] = scale * i;
    if (g_gpu) {
        const int n = N, one = 1;
        const double two = 2.0;
        g_dscal(&n, &two, v, &one);
        *expect_factor = 2.0 * scale;
    } else {
        if (g_set(v, DEVICE) != 0) check(0, "setup", "set_residency DEVICE");
        *expect_factor = scale;
    }
    return v;
}

/* a tracked vector whose view is read-only (SHARED): on the GPU after a
 * reading BLAS call uploaded it */
static double *shared_view(void)
{
    double *v = g_malloc(N * sizeof(double));
    for (int i = 0; i < N; ++i) v[i] = 1.0;
    if (g_gpu) {
        const int n = N, one = 1;
        (void)g_ddot(&n, v, &one, v, &one);
    } else if (g_set(v, SHARED) != 0) {
        check(0, "setup", "set_residency SHARED");
    }
    return v;
}

static int same(const double *a, double f)
{
    for (int i = 0; i < N; ++i)
        if (a[i] != f * i) return 0;
    return 1;
}

static int tmpfile_fd(const char *tag)
{
    char p[512];
    snprintf(p, sizeof p, "%s/%s", g_dir, tag);
    return open(p, O_RDWR | O_CREAT | O_TRUNC, 0600);
}

static double *plain_read_back(int fd)
{
    double *b = malloc(N * sizeof(double));
    if (pread(fd, b, N * sizeof(double), 0) != (ssize_t)(N * sizeof(double))) {
        free(b);
        return NULL;
    }
    return b;
}

static void case_write(void)
{
    double f;
    double *v = device_newer(3.0, &f);
    int fd = tmpfile_fd("w");
    const ssize_t w = write(fd, v, N * sizeof(double));
    check(w == (ssize_t)(N * sizeof(double)), "write_device_object", strerror(errno));
    double *b = plain_read_back(fd);
    check(b && same(b, f), "write_device_object_contents", "file differs from the device data");
    check(g_get(v) == SHARED, "write_leaves_shared", "state after write(2) is not SHARED");
    free(b);
    close(fd);
    g_free(v);
}

static void case_read(void)
{
    double *src = malloc(N * sizeof(double));
    for (int i = 0; i < N; ++i) src[i] = -1.0 * i;
    int fd = tmpfile_fd("r");
    if (write(fd, src, N * sizeof(double)) != (ssize_t)(N * sizeof(double))) check(0, "read", "setup");
    double *v = shared_view();
    const ssize_t r = pread(fd, v, N * sizeof(double), 0);
    check(r == (ssize_t)(N * sizeof(double)), "pread_into_shared_object", strerror(errno));
    check(same(v, -1.0), "pread_into_shared_object_contents", "buffer differs from the file");
    check(g_get(v) == HOST, "read_demotes_to_host", "state after read(2) is not HOST");
    if (g_gpu) {
        /* the next BLAS call sees what read(2) stored */
        const int n = N, one = 1;
        double want = 0;
        for (int i = 0; i < N; ++i) want += (double)i * i;
        check(g_ddot(&n, v, &one, v, &one) == want, "blas_after_read", "ddot of the read data");
    }
    close(fd);
    g_free(v);
    free(src);
}

static void case_vectored_and_stdio(void)
{
    double f1, f2;
    double *a = device_newer(1.0, &f1);
    double *b = device_newer(5.0, &f2);
    int fd = tmpfile_fd("v");
    struct iovec iov[2] = {{a, N * sizeof(double)}, {b, N * sizeof(double)}};
    const ssize_t w = writev(fd, iov, 2);
    check(w == (ssize_t)(2 * N * sizeof(double)), "writev_device_objects", strerror(errno));
    double *chk = malloc(2 * N * sizeof(double));
    int ok = pread(fd, chk, 2 * N * sizeof(double), 0) == (ssize_t)(2 * N * sizeof(double));
    ok = ok && same(chk, f1) && same(chk + N, f2);
    check(ok, "writev_contents", "file differs");
    /* readv into two read-only views */
    double *c = shared_view(), *d = shared_view();
    struct iovec in[2] = {{c, N * sizeof(double)}, {d, N * sizeof(double)}};
    const ssize_t r = preadv(fd, in, 2, 0);
    check(r == (ssize_t)(2 * N * sizeof(double)) && same(c, f1) && same(d, f2), "preadv_into_shared",
          strerror(errno));
    close(fd);
    /* stdio: fwrite of a device-newer object, fread into a read-only view */
    double f3;
    double *e = device_newer(7.0, &f3);
    char p[512];
    snprintf(p, sizeof p, "%s/s", g_dir);
    FILE *fp = fopen(p, "w+b");
    check(fp && fwrite(e, sizeof(double), N, fp) == N, "fwrite_device_object", strerror(errno));
    double *g = shared_view();
    rewind(fp);
    check(fread(g, sizeof(double), N, fp) == N && same(g, f3), "fread_into_shared", "contents");
    fclose(fp);
    free(chk);
    g_free(a);
    g_free(b);
    g_free(c);
    g_free(d);
    g_free(e);
    g_free(g);
}

/* ---- the application's own SIGSEGV handler, installed after the library's ---- */
static volatile int g_app_calls = 0;
static char *g_guard_page = NULL;
static void app_segv(int sig, siginfo_t *si, void *uc)
{
    (void)sig;
    (void)uc;
    g_app_calls++;
    /* the application's own lazily-mapped page: make it accessible */
    if ((char *)si->si_addr >= g_guard_page && (char *)si->si_addr < g_guard_page + 4096)
        mprotect(g_guard_page, 4096, PROT_READ | PROT_WRITE);
    else
        _exit(77);
}
static void app_segv_plain(int sig)
{
    (void)sig;
    g_app_calls++;
    mprotect(g_guard_page, 4096, PROT_READ | PROT_WRITE);
}

static void case_app_handler(void)
{
    struct sigaction sa, old;
    memset(&sa, 0, sizeof sa);
    sa.sa_sigaction = app_segv;
    sa.sa_flags = SA_SIGINFO;
    sigemptyset(&sa.sa_mask);
    check(sigaction(SIGSEGV, &sa, &old) == 0, "app_sigaction_installs", strerror(errno));
    check(!(old.sa_flags & SA_SIGINFO) && old.sa_handler == SIG_DFL, "app_sees_default_before",
          "the previous handler reported to the application is not SIG_DFL");
    /* a tracked object's fault stays the library's */
    double f;
    double *v = device_newer(2.0, &f);
    volatile double s = 0;
    for (int i = 0; i < N; i += 4096) s += v[i];
    check(g_app_calls == 0 && same(v, f), "tracked_fault_not_chained",
          "the application's handler saw a tracked-object fault");
    /* any other fault reaches the application's handler */
    g_guard_page = mmap(NULL, 4096, PROT_NONE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    g_guard_page[10] = 42;
    check(g_app_calls == 1 && g_guard_page[10] == 42, "foreign_fault_chained",
          "the application's handler did not run for its own fault");
    /* sigaction reports the application's handler back */
    struct sigaction cur;
    sigaction(SIGSEGV, NULL, &cur);
    check((cur.sa_flags & SA_SIGINFO) && cur.sa_sigaction == app_segv, "app_sees_its_handler",
          "sigaction(SIGSEGV, NULL, &old) does not return the application's handler");
    /* signal() replaces it the same way */
    mprotect(g_guard_page, 4096, PROT_NONE);
    signal(SIGSEGV, app_segv_plain);
    double *w = device_newer(4.0, &f);
    for (int i = 0; i < N; i += 4096) s += w[i];
    g_guard_page[11] = 43;
    check(g_app_calls == 2 && same(w, f), "signal_handler_chained", "signal() path");
    signal(SIGSEGV, SIG_DFL);
    munmap(g_guard_page, 4096);
    g_free(v);
    g_free(w);
}

/* ---- faults from several threads at once (served by the helper thread) ---- */
struct th {
    double *v;
    double f;
    int ok;
};
static void *touch(void *arg)
{
    struct th *t = arg;
    t->ok = same(t->v, t->f);
    return NULL;
}
static void case_threads(void)
{
    struct th t[8];
    pthread_t id[8];
    for (int i = 0; i < 8; ++i) t[i].v = device_newer(1.0 + i, &t[i].f);
    for (int i = 0; i < 8; ++i) pthread_create(&id[i], NULL, touch, &t[i]);
    int ok = 1;
    for (int i = 0; i < 8; ++i) {
        pthread_join(id[i], NULL);
        ok &= t[i].ok;
        g_free(t[i].v);
    }
    check(ok, "concurrent_faults", "a thread saw wrong data");
}

/* ---- fork: the child gets a snapshot, neither sees the other's writes ---- */
static void case_fork(void)
{
    double f;
    double *v = device_newer(6.0, &f);
    double *h = g_malloc(N * sizeof(double)); /* HOST state, writable */
    for (int i = 0; i < N; ++i) h[i] = 1.0 * i;
    pid_t pid = fork();
    if (pid == 0) {
        int bad = 0;
        bad |= !same(v, f);          /* the device data, downloaded at the fork */
        bad |= !same(h, 1.0) << 1;
        v[1] = -5.0;                 /* private to the child */
        h[2] = -7.0;
        bad |= (g_get(v) != -1) << 2; /* no longer tracked in the child */
        g_free(h);
        _exit(bad);
    }
    int st = 0;
    waitpid(pid, &st, 0);
    check(WIFEXITED(st) && WEXITSTATUS(st) == 0, "fork_child_sees_snapshot",
          "child exit status names the mismatch (bit 0 device data, 1 host data, 2 tracking)");
    check(same(v, f) && same(h, 1.0), "fork_parent_isolated", "the child's writes reached the parent");
    check(g_get(v) >= 0 && g_get(h) >= 0, "fork_parent_still_tracks", "parent lost its objects");
    if (g_gpu) {
        const int n = N, one = 1;
        double want = 0;
        for (int i = 0; i < N; ++i) want += (double)i * i;
        check(g_ddot(&n, h, &one, h, &one) == want, "blas_after_fork", "parent GPU use after fork");
    }
    h[3] = 3.0; /* the parent's view is writable again */
    g_free(v);
    g_free(h);
}

int main(int argc, char **argv)
{
    g_gpu = argc > 1 && !strcmp(argv[1], "gpu");
    snprintf(g_dir, sizeof g_dir, "%s", argc > 2 ? argv[2] : "/tmp");
    g_malloc = (void *(*)(size_t))sym("gpublas_malloc");
    g_free = (void (*)(void *))sym("gpublas_free");
    g_set = (int (*)(void *, int))sym("gpublas_debug_set_residency");
    g_get = (int (*)(const void *))sym("gpublas_debug_residency");
    if (g_gpu) {
        g_dscal = (void (*)(const int *, const double *, double *, const int *))sym("dscal_");
        g_ddot = (double (*)(const int *, const double *, const int *, const double *,
                             const int *))sym("ddot_");
    }
    case_write();
    case_read();
    case_vectored_and_stdio();
    case_app_handler();
    case_threads();
    case_fork();
    printf("failures %d\n", g_fail);
    return g_fail;
}
