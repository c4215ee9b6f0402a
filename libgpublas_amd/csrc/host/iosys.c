/*
 * iosys.c -- interposers that keep tracked memory transparent to system
 * calls, stdio and signal handling; linked into libgpublas_preload.so only.
 *
 * Reference counterpart: the managed memory of blas2cuda.c:127-148, whose
 * pages the kernel migrates for any access, system calls included.  Here a
 * device-newer tracked object is protected PROT_NONE and an application
 * access faults into the library's SIGSEGV handler -- but a system call that
 * touches such a page does not fault, it fails with EFAULT.  So every call
 * that hands user memory to the kernel, or (fread/fwrite) may do so behind a
 * FILE buffer, first brings the tracked objects it touches to the host:
 *   read, pread, readv, preadv, fread     store into the buffer (view opened
 *                                          read-write, object demoted to HOST)
 *   write, pwrite, writev, pwritev, fwrite read the buffer (device copy
 *                                          downloaded if newer, view read-only)
 * With no tracked object alive each interposer is one load and a tail call.
 *
 * sigaction / signal on SIGSEGV: the application's handler is recorded as
 * the one the library chains to for faults that are not its own, and the
 * library's handler stays installed.  fork: device-newer objects are
 * downloaded before the real fork, so the fork handlers (tracker.c) make no
 * GPU runtime call.
 */
#include "internal.h"
#include <dlfcn.h>
#include <errno.h>
#include <stdio.h>
#include <string.h>
#include <sys/uio.h>
#include <unistd.h>

void gb_fork_sync_device(void);

#define REAL(ret, name, args)                                                    \
    static ret(*real_##name) args = NULL;                                        \
    static ret(*get_##name(void)) args                                           \
    {                                                                            \
        if (!real_##name) {                                                      \
            gb_enter();                                                          \
            real_##name = (ret(*) args)gb_libc_sym(#name);                       \
            gb_leave();                                                          \
        }                                                                        \
        return real_##name;                                                      \
    }

static inline int io_tracked(void)
{
    return gb_tracked_count() && !gb_guard;
}

static void iov_access(const struct iovec *iov, int cnt, int write)
{
    for (int i = 0; iov && i < cnt; ++i) gb_host_access(iov[i].iov_base, iov[i].iov_len, write);
}

REAL(ssize_t, read, (int, void *, size_t))
ssize_t read(int fd, void *buf, size_t n)
{
    if (io_tracked()) gb_host_access(buf, n, 1);
    return get_read()(fd, buf, n);
}

REAL(ssize_t, write, (int, const void *, size_t))
ssize_t write(int fd, const void *buf, size_t n)
{
    if (io_tracked()) gb_host_access(buf, n, 0);
    return get_write()(fd, buf, n);
}

REAL(ssize_t, pread, (int, void *, size_t, off_t))
ssize_t pread(int fd, void *buf, size_t n, off_t off)
{
    if (io_tracked()) gb_host_access(buf, n, 1);
    return get_pread()(fd, buf, n, off);
}

REAL(ssize_t, pread64, (int, void *, size_t, off64_t))
ssize_t pread64(int fd, void *buf, size_t n, off64_t off)
{
    if (io_tracked()) gb_host_access(buf, n, 1);
    return get_pread64()(fd, buf, n, off);
}

REAL(ssize_t, pwrite, (int, const void *, size_t, off_t))
ssize_t pwrite(int fd, const void *buf, size_t n, off_t off)
{
    if (io_tracked()) gb_host_access(buf, n, 0);
    return get_pwrite()(fd, buf, n, off);
}

REAL(ssize_t, pwrite64, (int, const void *, size_t, off64_t))
ssize_t pwrite64(int fd, const void *buf, size_t n, off64_t off)
{
    if (io_tracked()) gb_host_access(buf, n, 0);
    return get_pwrite64()(fd, buf, n, off);
}

REAL(ssize_t, readv, (int, const struct iovec *, int))
ssize_t readv(int fd, const struct iovec *iov, int cnt)
{
    if (io_tracked()) iov_access(iov, cnt, 1);
    return get_readv()(fd, iov, cnt);
}

REAL(ssize_t, writev, (int, const struct iovec *, int))
ssize_t writev(int fd, const struct iovec *iov, int cnt)
{
    if (io_tracked()) iov_access(iov, cnt, 0);
    return get_writev()(fd, iov, cnt);
}

REAL(ssize_t, preadv, (int, const struct iovec *, int, off_t))
ssize_t preadv(int fd, const struct iovec *iov, int cnt, off_t off)
{
    if (io_tracked()) iov_access(iov, cnt, 1);
    return get_preadv()(fd, iov, cnt, off);
}

REAL(ssize_t, pwritev, (int, const struct iovec *, int, off_t))
ssize_t pwritev(int fd, const struct iovec *iov, int cnt, off_t off)
{
    if (io_tracked()) iov_access(iov, cnt, 0);
    return get_pwritev()(fd, iov, cnt, off);
}

REAL(size_t, fread, (void *, size_t, size_t, FILE *))
size_t fread(void *buf, size_t size, size_t nmemb, FILE *f)
{
    if (io_tracked() && size && nmemb <= (size_t)-1 / size) gb_host_access(buf, size * nmemb, 1);
    return get_fread()(buf, size, nmemb, f);
}

REAL(size_t, fwrite, (const void *, size_t, size_t, FILE *))
size_t fwrite(const void *buf, size_t size, size_t nmemb, FILE *f)
{
    if (io_tracked() && size && nmemb <= (size_t)-1 / size) gb_host_access(buf, size * nmemb, 0);
    return get_fwrite()(buf, size, nmemb, f);
}

/* ---- SIGSEGV handler chaining ---- */
int sigaction(int sig, const struct sigaction *act, struct sigaction *old)
{
    if (sig == SIGSEGV && gb_segv_app_action(act, old)) return 0;
    return gb_real_sigaction(sig, act, old);
}

REAL(sighandler_t, signal, (int, sighandler_t))
sighandler_t signal(int sig, sighandler_t h)
{
    if (sig == SIGSEGV && h != SIG_ERR) {
        struct sigaction sa, old;
        memset(&sa, 0, sizeof sa);
        sa.sa_handler = h;
        sa.sa_flags = SA_RESTART; /* glibc's signal(): BSD semantics */
        sigemptyset(&sa.sa_mask);
        if (gb_segv_app_action(&sa, &old))
            return (old.sa_flags & SA_SIGINFO) ? (sighandler_t)(void *)old.sa_sigaction
                                                : old.sa_handler;
    }
    return get_signal()(sig, h);
}

/* ---- fork ---- */
REAL(pid_t, fork, (void))
pid_t fork(void)
{
    gb_fork_sync_device();
    return get_fork()();
}
