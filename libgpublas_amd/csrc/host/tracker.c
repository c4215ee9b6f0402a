/*
 * tracker.c -- tracked objects: host memory whose contents stay resident in
 * HBM across BLAS calls, kept coherent with the host by page protection.
 *
 * Reference counterparts: the managed-memory object manager
 * (blas2cuda.c:127-175), the object table with interior-pointer lookup
 * (lib/obj_tracker.c:536-637, 669-787) and the hit/miss residency logic of
 * gpuptr (runtime-mem.hpp:37-165).
 *
 * Why not managed memory: the reference allocates tracked objects with
 * cudaMallocManaged and lets page migration keep them on the device.  On an
 * MI355X without XNACK (the configuration of these systems), hipMallocManaged
 * returns host memory that kernels read over the host link at tens of GB/s,
 * two orders of magnitude below HBM.  So a tracked object here is:
 *   - a memfd mapped twice: the application view (returned by malloc) and a
 *     private alias that the library always reads/writes;
 *   - a device mirror in HBM, allocated on first BLAS use;
 *   - a state: HOST (host copy newer, device stale or absent), SHARED (both
 *     equal, application view read-only) or DEVICE (device copy newer,
 *     application view inaccessible).
 * A BLAS call uploads a HOST object once and leaves it SHARED; a call that
 * writes it leaves it DEVICE.  A host access to a protected page raises
 * SIGSEGV; the handler downloads a DEVICE object into the alias view and
 * re-opens the application view (read-only for a read, read-write for a
 * write), so host code sees exactly what the GPU wrote, and a host write
 * demotes the object to HOST so the next BLAS call re-uploads it.  Objects
 * used only by the GPU (the common BLAS chain) never cross the host link
 * after their first upload.
 */
#include "internal.h"
#include <errno.h>
#include <signal.h>
#include <stdlib.h>
#include <string.h>
#include <ucontext.h>
#include <unistd.h>
#include <dlfcn.h>
#include <fcntl.h>
#include <linux/futex.h>
#include <poll.h>
#include <time.h>
#include <sys/mman.h>
#include <sys/syscall.h>

extern void *__libc_malloc(size_t);
extern void *__libc_calloc(size_t, size_t);
extern void __libc_free(void *);

/* ---- object table: sorted array of objects by application address ---- */
static pthread_rwlock_t g_tab_lock = PTHREAD_RWLOCK_INITIALIZER;
static struct gb_obj **g_tab = NULL;
static size_t g_tab_n = 0, g_tab_cap = 0;
static uint64_t g_uid = 0;
static size_t g_page = 0;

static size_t page_size(void)
{
    if (!g_page) g_page = (size_t)sysconf(_SC_PAGESIZE);
    return g_page;
}

/* index of the first object with host > p (upper bound) */
static size_t tab_upper(const char *p)
{
    size_t lo = 0, hi = g_tab_n;
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        if (g_tab[mid]->host <= p) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

static void tab_insert(struct gb_obj *o)
{
    pthread_rwlock_wrlock(&g_tab_lock);
    if (g_tab_n == g_tab_cap) {
        size_t nc = g_tab_cap ? g_tab_cap * 2 : 1024;
        void *nt = mmap(NULL, nc * sizeof(*g_tab), PROT_READ | PROT_WRITE,
                        MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
        if (nt == MAP_FAILED) gb_fatal("object table growth failed");
        if (g_tab) {
            memcpy(nt, g_tab, g_tab_n * sizeof(*g_tab));
            munmap(g_tab, g_tab_cap * sizeof(*g_tab));
        }
        g_tab = nt;
        g_tab_cap = nc;
    }
    size_t pos = tab_upper(o->host);
    memmove(&g_tab[pos + 1], &g_tab[pos], (g_tab_n - pos) * sizeof(*g_tab));
    g_tab[pos] = o;
    g_tab_n++;
    pthread_rwlock_unlock(&g_tab_lock);
}

static struct gb_obj *tab_remove(const void *p)
{
    struct gb_obj *o = NULL;
    pthread_rwlock_wrlock(&g_tab_lock);
    size_t pos = tab_upper(p);
    if (pos > 0 && g_tab[pos - 1]->host == (const char *)p) {
        o = g_tab[pos - 1];
        memmove(&g_tab[pos - 1], &g_tab[pos], (g_tab_n - pos) * sizeof(*g_tab));
        g_tab_n--;
    }
    pthread_rwlock_unlock(&g_tab_lock);
    return o;
}

/* object whose mapping contains p, detached or not */
static struct gb_obj *tab_find(const void *p)
{
    if (!__atomic_load_n(&g_tab_n, __ATOMIC_RELAXED) || !p) return NULL;
    struct gb_obj *o = NULL;
    pthread_rwlock_rdlock(&g_tab_lock);
    size_t pos = tab_upper(p);
    if (pos > 0) {
        struct gb_obj *c = g_tab[pos - 1];
        if ((const char *)p < c->host + c->map_size) o = c;
    }
    pthread_rwlock_unlock(&g_tab_lock);
    return o;
}

struct gb_obj *gb_obj_find(const void *p)
{
    struct gb_obj *o = tab_find(p);
    return o && !o->detached ? o : NULL;
}

int gb_obj_owns(const void *p)
{
    struct gb_obj *o = tab_find(p);
    return o && o->host == (const char *)p;
}

int gb_tracked_count(void) { return (int)g_tab_n; }

void gb_obj_foreach_used(void (*fn)(uint64_t nth, void *callsite, size_t size))
{
    pthread_rwlock_rdlock(&g_tab_lock);
    for (size_t k = 0; k < g_tab_n; ++k)
        if (g_tab[k]->used) fn(g_tab[k]->nth, g_tab[k]->callsite, g_tab[k]->size);
    pthread_rwlock_unlock(&g_tab_lock);
}
bool gb_ptr_tracked(const void *p) { return gb_obj_find(p) != NULL; }

size_t gb_obj_usable_size(const void *p)
{
    struct gb_obj *o = tab_find(p);
    if (!o || o->host != (const char *)p) return 0;
    return o->size;
}

/* ---- creation / destruction ---- */
void *gb_obj_create(size_t size, uint64_t nth, void *callsite, int zero)
{
    (void)zero; /* memfd pages start zeroed */
    gb_enter();
    const size_t ps = page_size();
    const size_t map = ((size ? size : 1) + ps - 1) / ps * ps;
    int fd = (int)syscall(SYS_memfd_create, "gpublas-object", 1u /* MFD_CLOEXEC */);
    if (fd < 0) {
        gb_leave();
        return NULL;
    }
    if (ftruncate(fd, (off_t)map) != 0) {
        close(fd);
        gb_leave();
        return NULL;
    }
    char *host = mmap(NULL, map, PROT_READ | PROT_WRITE, MAP_SHARED, fd, 0);
    char *alias = mmap(NULL, map, PROT_READ | PROT_WRITE, MAP_SHARED, fd, 0);
    close(fd);
    if (host == MAP_FAILED || alias == MAP_FAILED) {
        if (host != MAP_FAILED) munmap(host, map);
        if (alias != MAP_FAILED) munmap(alias, map);
        gb_leave();
        return NULL;
    }
    struct gb_obj *o = __libc_calloc(1, sizeof *o);
    if (!o) gb_fatal("out of host memory for object record");
    o->host = host;
    o->alias = alias;
    o->size = size;
    o->map_size = map;
    o->nth = nth;
    o->uid = __atomic_fetch_add(&g_uid, 1, __ATOMIC_RELAXED);
    o->callsite = callsite;
    o->state = ST_HOST;
    pthread_mutex_init(&o->mu, NULL);
    tab_insert(o);
    GB_COUNT(tracked_bytes, size);
    gb_trace_alloc('T', nth, callsite, size, host);
    gb_leave();
    return host;
}

size_t gb_resident_bytes = 0;
static uint64_t g_tick = 0;
static void fault_service_start(void);

static void wait_use(void *stream, uint64_t seq)
{
    gb_check(gpublas_dev_stream_wait(stream, seq), "wait for a tracked object");
}

/* Caller holds o->mu.  Wait (host-side) for the kernels on streams other
 * than `stream` that a use must follow: the last writer, and for a write
 * every recorded reader.  Entries found complete are dropped. */
static void wait_hazards_locked(struct gb_obj *o, void *stream, int write)
{
    if (o->wr_stream && o->wr_stream != stream) {
        wait_use(o->wr_stream, o->wr_seq);
        o->wr_stream = NULL;
    }
    if (write) {
        int k = 0;
        for (int i = 0; i < o->nrd; ++i) {
            if (o->rd_stream[i] != stream) {
                wait_use(o->rd_stream[i], o->rd_seq[i]);
                continue;
            }
            o->rd_stream[k] = o->rd_stream[i];
            o->rd_seq[k++] = o->rd_seq[i];
        }
        o->nrd = k;
    }
}

void gb_obj_quiesce(struct gb_obj *o) { wait_hazards_locked(o, NULL, 1); }

void gb_obj_wait_hazards(struct gb_obj *o, int write)
{
    pthread_mutex_lock(&o->mu);
    wait_hazards_locked(o, NULL, write);
    pthread_mutex_unlock(&o->mu);
}

void gb_obj_unref(struct gb_obj *o)
{
    pthread_mutex_lock(&o->mu);
    o->refs--;
    pthread_mutex_unlock(&o->mu);
}

static int idle_locked(struct gb_obj *o)
{
    if (o->refs || !o->dev || o->state != ST_SHARED) return 0;
    if (o->wr_stream && !gpublas_dev_stream_done(o->wr_stream, o->wr_seq)) return 0;
    for (int i = 0; i < o->nrd; ++i)
        if (!gpublas_dev_stream_done(o->rd_stream[i], o->rd_seq[i])) return 0;
    return 1;
}

static void drop_mirror_locked(struct gb_obj *o)
{
    gpublas_dev_release(o->dev);
    o->dev = NULL;
    o->wr_stream = NULL;
    o->nrd = 0;
    __atomic_fetch_sub(&gb_resident_bytes, o->map_size, __ATOMIC_RELAXED);
    /* the host copy is now the only one: let the application write it */
    mprotect(o->host, o->map_size, PROT_READ | PROT_WRITE);
    o->state = ST_HOST;
}

size_t gb_evict_idle(size_t want, struct gb_obj *const *keep, int nkeep)
{
    size_t freed = 0;
    while (freed < want) {
        struct gb_obj *best = NULL;
        pthread_rwlock_rdlock(&g_tab_lock);
        for (size_t i = 0; i < g_tab_n; ++i) {
            struct gb_obj *o = g_tab[i];
            int kept = 0;
            for (int j = 0; j < nkeep; ++j) kept |= keep[j] == o;
            if (kept || !o->dev || o->refs || o->state != ST_SHARED) continue;
            if (!best || o->last_use < best->last_use) best = o;
        }
        int done = 0;
        if (best) {
            pthread_mutex_lock(&best->mu);
            if (idle_locked(best)) {
                freed += best->map_size;
                drop_mirror_locked(best);
                GB_COUNT(evictions, 1);
            } else {
                /* busy after all: stop rather than spin on it */
                done = 1;
            }
            pthread_mutex_unlock(&best->mu);
        }
        pthread_rwlock_unlock(&g_tab_lock);
        if (!best || done) break;
    }
    return freed;
}

static void obj_free(struct gb_obj *o)
{
    gb_trace_alloc('U', o->nth, o->callsite, o->size, o->host);
    if (o->used && gb_opt.objtrack_out[0]) gb_objtrack_note(o->nth, o->callsite, o->size);
    if (!o->detached)
        gb_obj_quiesce(o); /* a kernel launched by a call that returned early may still use it */
    munmap(o->host, o->map_size);
    if (o->alias) munmap(o->alias, o->map_size);
    if (o->dev) {
        gpublas_dev_release(o->dev);
        __atomic_fetch_sub(&gb_resident_bytes, o->map_size, __ATOMIC_RELAXED);
    }
    __atomic_fetch_sub(&gb_stats.tracked_bytes, (uint64_t)o->size, __ATOMIC_RELAXED);
    pthread_mutex_destroy(&o->mu);
    __libc_free(o);
}

int gb_obj_destroy(void *p)
{
    if (!__atomic_load_n(&g_tab_n, __ATOMIC_RELAXED) || !p) return 0;
    gb_enter();
    struct gb_obj *o = tab_remove(p);
    if (o) obj_free(o);
    gb_leave();
    return o != NULL;
}

static void obj_pull_to_host(struct gb_obj *o)
{
    /* caller holds o->mu; an object without a mirror has no newer copy to fetch */
    if (o->state == ST_DEVICE && o->dev) {
        gb_obj_quiesce(o);
        void *s = gpublas_dev_stream();
        gb_check(gpublas_dev_d2h(o->alias, o->dev, o->size, s), "download of a tracked object");
        gb_check(gpublas_dev_sync(s), "download of a tracked object");
        GB_COUNT(d2h_bytes, o->size);
    }
}

void *gb_obj_realloc(void *p, size_t size)
{
    gb_enter();
    struct gb_obj *o = tab_find(p);
    if (!o || o->host != (char *)p) {
        gb_leave();
        return NULL;
    }
    if (o->detached) {
        /* a fork child's copy: becomes ordinary heap memory */
        void *np = __libc_malloc(size);
        if (np) {
            memcpy(np, o->host, size < o->size ? size : o->size);
            gb_obj_destroy(p);
        }
        gb_leave();
        return np;
    }
    void *np = gb_obj_create(size, o->nth, o->callsite, 0);
    if (np) {
        pthread_mutex_lock(&o->mu);
        obj_pull_to_host(o);
        struct gb_obj *n = gb_obj_find(np);
        memcpy(n->alias, o->alias, size < o->size ? size : o->size);
        pthread_mutex_unlock(&o->mu);
        gb_obj_destroy(p);
    }
    gb_leave();
    return np;
}

/* ---- coherence ---- */
int gb_obj_to_device(struct gb_obj *o, void *stream, int write)
{
    int uploaded = 0;
    pthread_mutex_lock(&o->mu);
    o->refs++;
    o->last_use = __atomic_add_fetch(&g_tick, 1, __ATOMIC_RELAXED);
    if (!o->dev) {
        fault_service_start(); /* host touches of device-newer objects are served there */
        o->dev = gpublas_dev_alloc(o->map_size);
        if (!o->dev) gb_fatal("HBM exhausted allocating a %zu-byte tracked object", o->map_size);
        __atomic_fetch_add(&gb_resident_bytes, o->map_size, __ATOMIC_RELAXED);
    }
    /* kernels queued on other streams may still use the mirror: an upload
     * or a write must follow all of them, a read only the last writer */
    wait_hazards_locked(o, stream, write || o->state == ST_HOST);
    if (o->state == ST_HOST) {
        uploaded = 1;
        /* protect first: a host write racing the upload faults and demotes */
        mprotect(o->host, o->map_size, PROT_READ);
        gb_check(gpublas_dev_h2d(o->dev, o->alias, o->size, stream), "upload of a tracked object");
        o->state = ST_SHARED;
        GB_COUNT(uploads, 1);
        GB_COUNT(h2d_bytes, o->size);
    }
    o->used = 1;
    pthread_mutex_unlock(&o->mu);
    return uploaded;
}

void gb_obj_device_written(struct gb_obj *o)
{
    pthread_mutex_lock(&o->mu);
    if (o->state != ST_DEVICE) {
        mprotect(o->host, o->map_size, PROT_NONE);
        o->state = ST_DEVICE;
    }
    pthread_mutex_unlock(&o->mu);
}

/* Caller holds o->mu: make the application view usable for a host access
 * (download a DEVICE object first).  A read leaves the object SHARED (view
 * read-only), a write HOST (view read-write; the next BLAS call uploads). */
static void host_open_locked(struct gb_obj *o, int write)
{
    if (o->state == ST_DEVICE) {
        obj_pull_to_host(o);
        if (write) {
            mprotect(o->host, o->map_size, PROT_READ | PROT_WRITE);
            o->state = ST_HOST;
        } else {
            mprotect(o->host, o->map_size, PROT_READ);
            o->state = ST_SHARED;
        }
    } else if (o->state == ST_SHARED && write) {
        mprotect(o->host, o->map_size, PROT_READ | PROT_WRITE);
        o->state = ST_HOST;
    }
}

static int handle_fault(void *addr, int write)
{
    struct gb_obj *o = gb_obj_find(addr);
    if (!o) return 0;
    gb_enter();
    pthread_mutex_lock(&o->mu);
    /* ST_HOST (or SHARED on a read): another thread already opened the view;
     * the access is retried */
    if (o->state == ST_DEVICE || (o->state == ST_SHARED && write)) GB_COUNT(faults, 1);
    host_open_locked(o, write);
    pthread_mutex_unlock(&o->mu);
    gb_leave();
    return 1;
}

void gb_host_access(const void *p, size_t n, int write)
{
    if (!n || !p || gb_guard || !__atomic_load_n(&g_tab_n, __ATOMIC_RELAXED)) return;
    const char *lo = p, *hi = lo + n < lo ? (const char *)UINTPTR_MAX : lo + n;
    gb_enter();
    /* objects overlapping [lo, hi), a batch at a time: the table lock is not
     * held across downloads (the runtime may allocate meanwhile) */
    for (;;) {
        struct gb_obj *batch[16];
        int nb = 0;
        pthread_rwlock_rdlock(&g_tab_lock);
        size_t k = tab_upper(lo);
        if (k > 0 && g_tab[k - 1]->host + g_tab[k - 1]->map_size > lo) --k;
        for (; k < g_tab_n && g_tab[k]->host < hi && nb < 16; ++k)
            if (!g_tab[k]->detached) batch[nb++] = g_tab[k];
        const char *next = k < g_tab_n ? g_tab[k]->host : hi;
        pthread_rwlock_unlock(&g_tab_lock);
        for (int i = 0; i < nb; ++i) {
            pthread_mutex_lock(&batch[i]->mu);
            if (batch[i]->state == ST_DEVICE || (batch[i]->state == ST_SHARED && write))
                GB_COUNT(io_syncs, 1);
            host_open_locked(batch[i], write);
            pthread_mutex_unlock(&batch[i]->mu);
        }
        if (nb < 16 || next >= hi) break;
        lo = next;
    }
    gb_leave();
}

/* ---- fault service -------------------------------------------------------
 * The SIGSEGV handler does not download in signal context: it hands the
 * fault to a helper thread (a request on its own stack, pushed on a
 * lock-free list) and sleeps on a futex until the helper has opened the
 * view.  The GPU runtime, locks and allocation then run on an ordinary
 * thread.  Before the helper exists (no object has reached the device yet)
 * and on the helper itself the fault is handled in place. */
struct fault_req {
    void *addr;
    int write;
    int done;     /* futex word: 1 once handled */
    int handled;
    struct fault_req *next;
};
static struct fault_req *g_fault_head = NULL;
static int g_fault_bell = 0;       /* futex word: bumped per request */
static pthread_t g_fault_thread;
static int g_fault_up = 0;         /* 1 while the helper runs (0 again in a fork child) */
static long g_fault_pid = 0;       /* the process the helper runs in */
static pthread_mutex_t g_fault_start_mu = PTHREAD_MUTEX_INITIALIZER;

static long futex_op(int *w, int op, int val)
{
    return syscall(SYS_futex, w, op, val, NULL, NULL, 0);
}

static void *fault_service(void *arg)
{
    (void)arg;
    sigset_t all;
    sigfillset(&all);
    pthread_sigmask(SIG_BLOCK, &all, NULL);
    gb_enter(); /* nothing this thread allocates is tracked */
    for (;;) {
        const int bell = __atomic_load_n(&g_fault_bell, __ATOMIC_ACQUIRE);
        struct fault_req *r = __atomic_exchange_n(&g_fault_head, NULL, __ATOMIC_ACQ_REL);
        if (!r) {
            futex_op(&g_fault_bell, FUTEX_WAIT_PRIVATE, bell);
            continue;
        }
        while (r) {
            struct fault_req *next = r->next; /* r lives on the requester's stack */
            r->handled = handle_fault(r->addr, r->write);
            __atomic_store_n(&r->done, 1, __ATOMIC_RELEASE);
            futex_op(&r->done, FUTEX_WAKE_PRIVATE, 1);
            r = next;
        }
    }
    return NULL;
}

static void fault_service_start(void)
{
    if (__atomic_load_n(&g_fault_up, __ATOMIC_ACQUIRE)) return;
    pthread_mutex_lock(&g_fault_start_mu);
    if (!g_fault_up) {
        pthread_attr_t at;
        pthread_attr_init(&at);
        pthread_attr_setstacksize(&at, 256 << 10);
        g_fault_pid = syscall(SYS_getpid);
        if (pthread_create(&g_fault_thread, &at, fault_service, NULL) == 0)
            __atomic_store_n(&g_fault_up, 1, __ATOMIC_RELEASE);
        pthread_attr_destroy(&at);
    }
    pthread_mutex_unlock(&g_fault_start_mu);
}

/* signal context: async-signal-safe operations only (atomics, futex) */
static int request_fault(void *addr, int write)
{
    /* no helper yet, this is the helper, or this is a fork child whose fault
     * comes before the library's child handler ran (the helper is the
     * parent's): handle it here */
    if (!__atomic_load_n(&g_fault_up, __ATOMIC_ACQUIRE) ||
        syscall(SYS_getpid) != g_fault_pid || pthread_equal(pthread_self(), g_fault_thread))
        return handle_fault(addr, write);
    struct fault_req r = {addr, write, 0, 0, NULL};
    struct fault_req *h = __atomic_load_n(&g_fault_head, __ATOMIC_RELAXED);
    do {
        r.next = h;
    } while (!__atomic_compare_exchange_n(&g_fault_head, &h, &r, 1, __ATOMIC_RELEASE,
                                          __ATOMIC_RELAXED));
    __atomic_add_fetch(&g_fault_bell, 1, __ATOMIC_RELEASE);
    futex_op(&g_fault_bell, FUTEX_WAKE_PRIVATE, 1);
    while (!__atomic_load_n(&r.done, __ATOMIC_ACQUIRE)) futex_op(&r.done, FUTEX_WAIT_PRIVATE, 0);
    return r.handled;
}

/* ---- the SIGSEGV handler and the application's own ---------------------- */
static struct sigaction g_app_segv; /* what the application installed (or had before us) */
static int g_handler_installed = 0;
int gb_forked_child = 0;

/* libc's own definitions, whatever interposes them (the preload library's
 * sigaction, read, ...; RTLD_NEXT would miss libc when this library is
 * itself loaded with dlopen) */
void *gb_libc_sym(const char *name)
{
    static void *libc = NULL;
    if (!libc) {
        gb_enter();
        libc = dlopen("libc.so.6", RTLD_LAZY | RTLD_NOLOAD);
        gb_leave();
        if (!libc) return NULL;
    }
    return dlsym(libc, name);
}

extern int __sigaction(int, const struct sigaction *, struct sigaction *);

int gb_real_sigaction(int sig, const struct sigaction *act, struct sigaction *old)
{
    return __sigaction(sig, act, old);
}

static void on_segv(int sig, siginfo_t *si, void *ucv)
{
    const int saved_errno = errno;
    if (si && si->si_code == SEGV_ACCERR) {
        int write = 1;
#if defined(__x86_64__)
        const ucontext_t *uc = (const ucontext_t *)ucv;
        write = (uc->uc_mcontext.gregs[REG_ERR] & 2) != 0;
#endif
        if (request_fault(si->si_addr, write)) {
            errno = saved_errno;
            return;
        }
    }
    errno = saved_errno;
    /* not a tracked object: the application's handler, if any */
    const struct sigaction app = g_app_segv;
    if (app.sa_flags & SA_SIGINFO) {
        if (app.sa_sigaction) {
            app.sa_sigaction(sig, si, ucv);
            return;
        }
    } else if (app.sa_handler != SIG_DFL && app.sa_handler != SIG_IGN) {
        app.sa_handler(sig);
        return;
    }
    /* default action: re-executing the access faults again with SIG_DFL */
    struct sigaction dfl;
    memset(&dfl, 0, sizeof dfl);
    dfl.sa_handler = SIG_DFL;
    gb_real_sigaction(SIGSEGV, &dfl, NULL);
}

static int install_ours(struct sigaction *prev)
{
    struct sigaction sa;
    memset(&sa, 0, sizeof sa);
    sa.sa_sigaction = on_segv;
    sa.sa_flags = SA_SIGINFO | SA_NODEFER | SA_RESTART | SA_ONSTACK;
    sigemptyset(&sa.sa_mask);
    return gb_real_sigaction(SIGSEGV, &sa, prev);
}

static void fork_prepare(void);
static void fork_parent(void);
static void fork_child(void);

void gb_install_fault_handler(void)
{
    if (g_handler_installed) return;
    if (install_ours(&g_app_segv) == 0) g_handler_installed = 1;
    pthread_atfork(fork_prepare, fork_parent, fork_child);
}

int gb_segv_app_action(const struct sigaction *act, struct sigaction *old)
{
    if (!__atomic_load_n(&g_handler_installed, __ATOMIC_ACQUIRE)) return 0;
    struct sigaction prev = g_app_segv;
    if (act) g_app_segv = *act;
    if (old) *old = prev;
    return 1;
}

void gb_check_fault_handler(void)
{
    static uint64_t last_ns = 0;
    if (!g_handler_installed) return;
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC_COARSE, &t);
    const uint64_t now = (uint64_t)t.tv_sec * 1000000000ull + (uint64_t)t.tv_nsec;
    if (now - __atomic_load_n(&last_ns, __ATOMIC_RELAXED) < 1000000) return;
    __atomic_store_n(&last_ns, now, __ATOMIC_RELAXED);
    struct sigaction cur;
    if (gb_real_sigaction(SIGSEGV, NULL, &cur) != 0) return;
    if ((cur.sa_flags & SA_SIGINFO) && cur.sa_sigaction == on_segv) return;
    /* replaced behind the library's back (a raw syscall, another library's
     * static sigaction): chain to the newcomer and take the signal back */
    g_app_segv = cur;
    install_ours(NULL);
    GB_COUNT(handler_reinstalls, 1);
}

/* ---- fork -----------------------------------------------------------------
 * A child must not share the parent's tracked pages (a memfd mapped shared
 * would make the child's writes visible to the parent and the reverse).
 * Device-newer objects are downloaded before the fork (by the preload
 * library's fork(), else in the prepare handler).  The child copies every
 * object into private memory at the same address and detaches it: it is then
 * ordinary memory that free() releases, and every BLAS call of the child
 * runs on the host BLAS (the GPU runtime does not survive fork).  The
 * forking thread waits in the parent until the child has copied (bounded:
 * a child that dies first closes the pipe, one that stalls is given up on
 * with a warning), so the parent's own writes after fork() returns never
 * reach the child.  No lock of the library is held across the fork: other
 * libraries' fork handlers run in between, and threads they depend on may
 * be inside the library's interposers. */
static int g_fork_pipe[2] = {-1, -1};

void gb_fork_sync_device(void)
{
    if (!__atomic_load_n(&g_tab_n, __ATOMIC_RELAXED) || gb_forked_child) return;
    gb_enter();
    for (size_t k = 0;; ++k) {
        pthread_rwlock_rdlock(&g_tab_lock);
        struct gb_obj *o = k < g_tab_n ? g_tab[k] : NULL;
        pthread_rwlock_unlock(&g_tab_lock);
        if (!o) break;
        pthread_mutex_lock(&o->mu);
        if (o->state == ST_DEVICE && o->dev) host_open_locked(o, 0);
        pthread_mutex_unlock(&o->mu);
    }
    gb_leave();
}

static void fork_prepare(void)
{
    if (!__atomic_load_n(&g_tab_n, __ATOMIC_RELAXED) || gb_forked_child) return;
    gb_fork_sync_device();
    if (syscall(SYS_pipe2, g_fork_pipe, O_CLOEXEC) != 0) g_fork_pipe[0] = g_fork_pipe[1] = -1;
}

static void fork_parent(void)
{
    if (g_fork_pipe[0] < 0) return;
    syscall(SYS_close, g_fork_pipe[1]);
    /* the child writes one byte once it has copied (EOF if it died first);
     * allow 10 s plus 1 s per GB of tracked memory */
    const long budget_ms =
        10000 + (long)(__atomic_load_n(&gb_stats.tracked_bytes, __ATOMIC_RELAXED) >> 20);
    struct pollfd pf = {g_fork_pipe[0], POLLIN, 0};
    long waited = 0;
    int ready = 0;
    while (waited < budget_ms) {
        const int r = (int)syscall(SYS_poll, &pf, 1, 100);
        if (r > 0) {
            ready = 1;
            break;
        }
        if (r < 0 && errno != EINTR) break;
        waited += 100;
    }
    if (ready) {
        char c;
        (void)syscall(SYS_read, g_fork_pipe[0], &c, 1);
    } else {
        gb_log(2, "libgpublas: fork: the child did not finish copying its tracked objects in "
                  "time; continuing (it may see later writes of this process)\n");
    }
    syscall(SYS_close, g_fork_pipe[0]);
    g_fork_pipe[0] = g_fork_pipe[1] = -1;
}

static void fork_child(void)
{
    gb_forked_child = 1;
    gb_tracking = 0;
    gpublas_dev_forked_child();
    __atomic_store_n(&g_fault_up, 0, __ATOMIC_RELEASE);
    g_fault_head = NULL;
    /* locks possibly held by other threads of the parent at the fork */
    pthread_mutex_init(&g_fault_start_mu, NULL);
    pthread_rwlock_init(&g_tab_lock, NULL);
    for (size_t k = 0; k < g_tab_n; ++k) {
        struct gb_obj *o = g_tab[k];
        pthread_mutex_init(&o->mu, NULL);
        if (o->detached) continue;
        char *copy = mmap(NULL, o->map_size, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS,
                          -1, 0);
        if (copy == MAP_FAILED) gb_fatal("fork: no memory to copy a tracked object");
        memcpy(copy, o->alias, o->map_size);
        if (mremap(copy, o->map_size, o->map_size, MREMAP_MAYMOVE | MREMAP_FIXED, o->host) ==
            MAP_FAILED)
            gb_fatal("fork: cannot replace a tracked mapping");
        munmap(o->alias, o->map_size);
        o->alias = NULL;
        o->dev = NULL; /* the parent's HBM, not this process's */
        o->wr_stream = NULL;
        o->nrd = 0;
        o->refs = 0;
        o->state = ST_HOST;
        o->detached = 1;
    }
    if (g_fork_pipe[1] >= 0) {
        const char c = 1;
        syscall(SYS_write, g_fork_pipe[1], &c, 1);
        syscall(SYS_close, g_fork_pipe[1]);
        syscall(SYS_close, g_fork_pipe[0]);
        g_fork_pipe[0] = g_fork_pipe[1] = -1;
    }
}

/* ---- test hook ---- */
int gpublas_debug_set_residency(void *p, int state)
{
    struct gb_obj *o = gb_obj_find(p);
    if (!o || o->host != (char *)p || state < 0 || state > 2) return -1;
    int rc = 0;
    gb_enter();
    fault_service_start();
    pthread_mutex_lock(&o->mu);
    if (o->dev) {
        rc = -1; /* only for objects that never reached the device */
    } else {
        static const int prot[3] = {PROT_READ | PROT_WRITE, PROT_READ, PROT_NONE};
        static const int st[3] = {ST_HOST, ST_SHARED, ST_DEVICE};
        mprotect(o->host, o->map_size, prot[state]);
        o->state = st[state];
    }
    pthread_mutex_unlock(&o->mu);
    gb_leave();
    return rc;
}

int gpublas_debug_residency(const void *p)
{
    struct gb_obj *o = gb_obj_find(p);
    if (!o) return -1;
    pthread_mutex_lock(&o->mu);
    const int st = o->state == ST_HOST ? 0 : o->state == ST_SHARED ? 1 : 2;
    pthread_mutex_unlock(&o->mu);
    return st;
}

/* ---- operand residency ---- */
void gb_acquire_matrix(gb_operand *op, const void *host, size_t ld, size_t rows, size_t cols,
                       size_t es, int mode, void *stream)
{
    memset(op, 0, sizeof *op);
    op->host = host;
    op->ld = ld;
    op->rows = rows;
    op->cols = cols;
    op->es = es;
    op->mode = mode;
    if (!rows || !cols || !host) return;
    const size_t span = ((cols - 1) * ld + rows) * es;
    struct gb_obj *o = gb_obj_find(host);
    if (o && (const char *)host + span <= o->host + o->size) {
        op->uploaded = gb_obj_to_device(o, stream, (mode & GB_OUT) != 0);
        op->obj = o;
        op->dev = o->dev + ((const char *)host - o->host);
        gb_trace_call("operand", host);
        GB_COUNT(hits, 1);
        return;
    }
    op->tmp = gpublas_dev_alloc(span);
    if (!op->tmp) gb_fatal("HBM exhausted staging a %zu-byte operand", span);
    op->dev = op->tmp;
    if (mode & GB_IN) {
        int e = gpublas_dev_h2d_staged(op->dev, ld, host, ld, rows, cols, es, stream);
        gb_check(e, "operand upload");
        GB_COUNT(h2d_bytes, rows * cols * es);
        op->uploaded = 1;
    }
    GB_COUNT(misses, 1);
}

void gb_acquire_vector(gb_operand *op, const void *host, int n, int inc, size_t es, int mode,
                       void *stream)
{
    if (n <= 0) {
        memset(op, 0, sizeof *op);
        return;
    }
    const size_t ainc = (size_t)(inc < 0 ? -inc : inc);
    const size_t span = 1 + (size_t)(n - 1) * (ainc ? ainc : 1);
    /* a strided output has gaps that are copied back: bring them in too */
    if ((mode & GB_OUT) && ainc != 1) mode |= GB_IN;
    gb_acquire_matrix(op, host, span, span, 1, es, mode, stream);
}

void gb_acquire_packed(gb_operand *op, const void *host, int n, size_t es, int mode, void *stream)
{
    const size_t len = gb_packed_len(n);
    if (!len) {
        memset(op, 0, sizeof *op);
        return;
    }
    gb_acquire_matrix(op, host, len, len, 1, es, mode, stream);
}

void gb_release(gb_operand *op, void *stream)
{
    if (!op->dev) return;
    if (op->obj) {
        if (op->mode & GB_OUT) gb_obj_device_written(op->obj);
        return;
    }
    if (op->mode & GB_OUT) {
        int e = gpublas_dev_d2h_staged((void *)op->host, op->ld, op->dev, op->ld, op->rows,
                                       op->cols, op->es, stream);
        gb_check(e, "operand download");
        GB_COUNT(d2h_bytes, op->rows * op->cols * op->es);
    }
}

/* Record the call's kernels as the objects' latest writer / readers and
 * drop the references taken at acquire. */
static uint64_t mark_uses(void *stream, gb_operand *ops, int nops)
{
    const uint64_t seq = gpublas_dev_stream_mark(stream);
    for (int i = 0; i < nops; ++i) {
        struct gb_obj *o = ops[i].obj;
        if (!o) continue;
        pthread_mutex_lock(&o->mu);
        if ((ops[i].mode & GB_OUT) || ops[i].uploaded) {
            /* a write or this call's upload of the mirror (queued on `stream`):
             * readers on other streams were waited for at acquire */
            o->wr_stream = stream;
            o->wr_seq = seq;
            o->nrd = 0;
        } else {
            int j = 0;
            while (j < o->nrd && o->rd_stream[j] != stream) ++j;
            if (j == o->nrd) {
                if (o->nrd == GB_RD_SLOTS) {
                    /* table full: retire the oldest reader */
                    wait_use(o->rd_stream[0], o->rd_seq[0]);
                    for (int q = 1; q < o->nrd; ++q) {
                        o->rd_stream[q - 1] = o->rd_stream[q];
                        o->rd_seq[q - 1] = o->rd_seq[q];
                    }
                    j = --o->nrd;
                }
                o->rd_stream[j] = stream;
                o->nrd++;
            }
            o->rd_seq[j] = seq;
        }
        o->refs--;
        pthread_mutex_unlock(&o->mu);
    }
    return seq;
}

void gb_finish(void *stream, gb_operand *ops, int nops)
{
    int async_ok = 1;
    for (int i = 0; i < nops; ++i) {
        gb_release(&ops[i], stream);
        if (ops[i].dev && (!ops[i].obj || ops[i].uploaded)) async_ok = 0;
    }
    mark_uses(stream, ops, nops);
    if (async_ok) {
        GB_COUNT(async_returns, 1);
        return;
    }
    gb_check(gpublas_dev_sync(stream), "BLAS kernel");
    for (int i = 0; i < nops; ++i)
        if (ops[i].tmp) gpublas_dev_release(ops[i].tmp);
}

void gb_finish_scalar(void *stream, gb_operand *ops, int nops, uint64_t seq)
{
    const uint64_t mark = mark_uses(stream, ops, nops);
    gb_check(gpublas_dev_scalar_wait(stream, seq), "BLAS reduction kernel");
    /* the reduction has published, so the stream ran everything before it and
     * the reduction itself has read its operands */
    gpublas_dev_stream_observed(stream, mark);
    /* the kernel has consumed its inputs: staging blocks can be reused */
    for (int i = 0; i < nops; ++i)
        if (ops[i].tmp) gpublas_dev_release(ops[i].tmp);
}

bool gb_writers_done(const gb_operand *ops, int nops)
{
    for (int i = 0; i < nops; ++i) {
        struct gb_obj *o = ops[i].obj;
        if (!o || ops[i].uploaded) return false;  /* this call's upload is still queued */
        pthread_mutex_lock(&o->mu);
        /* known complete only (no stream query: that costs more than it saves) */
        const bool ok = !o->wr_stream || gpublas_dev_stream_known_done(o->wr_stream, o->wr_seq);
        pthread_mutex_unlock(&o->mu);
        if (!ok) return false;
    }
    return true;
}

void gb_finish_served(void *stream, gb_operand *ops, int nops)
{
    /* the server has read the operands and the caller has its result:
     * record the uses (ordering later writers on other streams) and drop
     * the references taken at acquire */
    mark_uses(stream, ops, nops);
    for (int i = 0; i < nops; ++i)
        if (ops[i].tmp) gpublas_dev_release(ops[i].tmp);
}

bool gb_any_tracked(const gb_operand *ops, int nops)
{
    for (int i = 0; i < nops; ++i)
        if (ops[i].obj) return true;
    return false;
}

/* ---- explicit API ---- */
void *gpublas_malloc(size_t bytes)
{
    return gb_obj_create(bytes, UINT64_MAX, __builtin_return_address(0), 0);
}

void *gpublas_calloc(size_t nmemb, size_t size)
{
    if (size && nmemb > (size_t)-1 / size) return NULL;
    return gb_obj_create(nmemb * size, UINT64_MAX, __builtin_return_address(0), 1);
}

void gpublas_free(void *p) { gb_obj_destroy(p); }

int gpublas_is_tracked(const void *p) { return gb_obj_find(p) != NULL; }

void *gpublas_device_ptr(const void *p)
{
    struct gb_obj *o = gb_obj_find(p);
    if (!o) return NULL;
    gb_enter();
    gb_require_device();
    void *s = gpublas_dev_stream();
    gb_obj_to_device(o, s, 1);
    gb_check(gpublas_dev_sync(s), "upload");
    gb_obj_unref(o);
    /* the caller may write through the device address: treat the device
     * copy as the newer one from now on */
    gb_obj_device_written(o);
    gb_leave();
    return o->dev + ((const char *)p - o->host);
}
