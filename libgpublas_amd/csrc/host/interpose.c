/*
 * interpose.c -- malloc / calloc / realloc / free / malloc_usable_size
 * interposers; linked into libgpublas_preload.so only.
 *
 * Reference counterpart: lib/obj_tracker.c:789-982.  An allocation made
 * outside the library (re-entrancy guard clear) gets the next ordinal; if
 * the heuristic selects it (ordinal, call site, size), it becomes a tracked
 * object (tracker.c), else it is served by glibc.  Creating a tracked object
 * never calls into the GPU runtime -- its device mirror is allocated at its
 * first BLAS use -- so these functions are safe to reach from any thread,
 * including the GPU runtime's own.
 */
#include "internal.h"
#include <dlfcn.h>
#include <errno.h>
#include <string.h>

extern void *__libc_malloc(size_t);
extern void *__libc_calloc(size_t, size_t);
extern void *__libc_realloc(void *, size_t);
extern void __libc_free(void *);

static inline int counting(void) { return gb_tracking && !gb_guard; }

static void *maybe_track(size_t n, void *ra, int zero)
{
    uint64_t nth = __atomic_fetch_add(&gb_nallocs, 1, __ATOMIC_RELAXED);
    if (!n || gb_opt.heuristic == H_FALSE) return NULL;
    gb_enter();
    bool t = gb_should_track(nth, ra, n);
    gb_leave();
    if (!t) return NULL;
    return gb_obj_create(n, nth, ra, zero);
}

void *malloc(size_t n)
{
    if (counting()) {
        void *p = maybe_track(n, __builtin_return_address(0), 0);
        if (p) return p;
    }
    return __libc_malloc(n);
}

void *calloc(size_t nmemb, size_t size)
{
    if (size && nmemb > (size_t)-1 / size) {
        errno = ENOMEM;
        return NULL;
    }
    if (counting()) {
        void *p = maybe_track(nmemb * size, __builtin_return_address(0), 1);
        if (p) return p;
    }
    return __libc_calloc(nmemb, size);
}

void *realloc(void *p, size_t n)
{
    if (!p) {
        if (counting()) {
            void *q = maybe_track(n, __builtin_return_address(0), 0);
            if (q) return q;
        }
        return __libc_realloc(NULL, n);
    }
    if (gb_tracked_count() && gb_obj_owns(p)) {
        if (!gb_guard) __atomic_fetch_add(&gb_nallocs, 1, __ATOMIC_RELAXED);
        if (n == 0) {
            gb_obj_destroy(p);
            return NULL;
        }
        return gb_obj_realloc(p, n);
    }
    if (counting()) __atomic_fetch_add(&gb_nallocs, 1, __ATOMIC_RELAXED);
    return __libc_realloc(p, n);
}

void free(void *p)
{
    if (!p) return;
    if (gb_tracked_count() && gb_obj_destroy(p)) return;
    __libc_free(p);
}

size_t malloc_usable_size(void *p)
{
    static size_t (*real)(void *) = NULL;
    if (!p) return 0;
    size_t s = gb_obj_usable_size(p);
    if (s) return s;
    if (!real) {
        gb_enter();
        real = (size_t(*)(void *))dlsym(RTLD_NEXT, "malloc_usable_size");
        gb_leave();
    }
    return real ? real(p) : 0;
}
