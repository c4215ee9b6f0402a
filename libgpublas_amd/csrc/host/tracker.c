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
#include <sys/mman.h>
#include <sys/syscall.h>

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

struct gb_obj *gb_obj_find(const void *p)
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
    struct gb_obj *o = gb_obj_find(p);
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
    gb_obj_quiesce(o); /* a kernel launched by a call that returned early may still use it */
    munmap(o->host, o->map_size);
    munmap(o->alias, o->map_size);
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
    /* caller holds o->mu */
    if (o->state == ST_DEVICE) {
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
    struct gb_obj *o = gb_obj_find(p);
    if (!o || o->host != (char *)p) {
        gb_leave();
        return NULL;
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

static int handle_fault(void *addr, int write)
{
    struct gb_obj *o = gb_obj_find(addr);
    if (!o) return 0;
    gb_enter();
    pthread_mutex_lock(&o->mu);
    if (o->state == ST_DEVICE) {
        obj_pull_to_host(o);
        if (write) {
            mprotect(o->host, o->map_size, PROT_READ | PROT_WRITE);
            o->state = ST_HOST;
        } else {
            mprotect(o->host, o->map_size, PROT_READ);
            o->state = ST_SHARED;
        }
        GB_COUNT(faults, 1);
    } else if (o->state == ST_SHARED) {
        mprotect(o->host, o->map_size, PROT_READ | PROT_WRITE);
        o->state = ST_HOST;
        GB_COUNT(faults, 1);
    }
    /* ST_HOST: another thread already opened the view; retry the access */
    pthread_mutex_unlock(&o->mu);
    gb_leave();
    return 1;
}

static struct sigaction g_old_segv;
static int g_handler_installed = 0;

static void on_segv(int sig, siginfo_t *si, void *ucv)
{
    if (si && si->si_code == SEGV_ACCERR) {
        int write = 1;
#if defined(__x86_64__)
        const ucontext_t *uc = (const ucontext_t *)ucv;
        write = (uc->uc_mcontext.gregs[REG_ERR] & 2) != 0;
#endif
        if (handle_fault(si->si_addr, write)) return;
    }
    if (g_old_segv.sa_flags & SA_SIGINFO) {
        if (g_old_segv.sa_sigaction) {
            g_old_segv.sa_sigaction(sig, si, ucv);
            return;
        }
    } else if (g_old_segv.sa_handler != SIG_DFL && g_old_segv.sa_handler != SIG_IGN) {
        g_old_segv.sa_handler(sig);
        return;
    }
    /* default action: re-executing the access faults again with SIG_DFL */
    struct sigaction dfl;
    memset(&dfl, 0, sizeof dfl);
    dfl.sa_handler = SIG_DFL;
    sigaction(SIGSEGV, &dfl, NULL);
}

void gb_install_fault_handler(void)
{
    if (g_handler_installed) return;
    struct sigaction sa;
    memset(&sa, 0, sizeof sa);
    sa.sa_sigaction = on_segv;
    sa.sa_flags = SA_SIGINFO | SA_NODEFER | SA_RESTART;
    sigemptyset(&sa.sa_mask);
    if (sigaction(SIGSEGV, &sa, &g_old_segv) == 0) g_handler_installed = 1;
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
