/*
 * objtrack.c -- which allocations become tracked objects, and the
 * object-tracking file that lists them.
 *
 * Reference counterparts: lib/oracle.c:14-73 (file loader and per-ordinal
 * bitmap), lib/obj_tracker.c:226-242 (heuristic switch) and the trace records
 * of the standalone tracer (lib/obj_tracker.c:426-483).  The capture at the
 * end (objtrack_out=<file>) writes the file directly from a run, which the
 * reference leaves to scripts/objtrace.sh:22-31 and scripts/analyze_trace.py.
 *
 * File format, one decision per line (the reference's "%1[HD] #%lu" format,
 * oracle.c:41, plus call-site keys):
 *     D #<nth> [anything]           allocation ordinal nth -> device-resident
 *     H #<nth> [anything]           allocation ordinal nth -> host
 *     D @<module>+0x<offset> [...]  every allocation made from that return
 *                                    address (module basename + offset from
 *                                    its load base) -> device-resident
 * Lines that match neither are ignored.  Ordinals count malloc / calloc /
 * realloc calls made outside the library, starting at 0; the ordinal of a
 * call is the counter value before the call (the reference queries one past
 * it, obj_tracker.c:824, and drops the last ordinal, oracle.c:22 -- both
 * corrected here).  Call-site keys do not depend on what other threads or
 * libraries allocate, so they survive runtime and library changes that shift
 * ordinals.
 */
#include "internal.h"
#include <dlfcn.h>
#include <errno.h>
#include <fcntl.h>
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#include <sys/mman.h>

uint64_t gb_nallocs = 0;
int gb_tracking = 0;

static uint8_t *g_bits = NULL;
static uint64_t g_nbits = 0;

struct cs_rule {
    char module[128];
    uint64_t offset;
    int device;
};
static struct cs_rule *g_rules = NULL;
static int g_nrules = 0, g_rules_cap = 0;

extern char *program_invocation_short_name;

static void *map_zeroed(size_t bytes)
{
    void *p = mmap(NULL, bytes, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    return p == MAP_FAILED ? NULL : p;
}

static void set_bit(uint64_t nth, int dev)
{
    if (nth >= g_nbits) {
        uint64_t nb = g_nbits ? g_nbits : 8192;
        while (nb <= nth) nb *= 2;
        uint8_t *n = map_zeroed(nb / 8);
        if (!n) return;
        if (g_bits) {
            memcpy(n, g_bits, g_nbits / 8);
            munmap(g_bits, g_nbits / 8);
        }
        g_bits = n;
        g_nbits = nb;
    }
    if (dev) g_bits[nth / 8] |= (uint8_t)(1u << (nth % 8));
    else g_bits[nth / 8] &= (uint8_t)~(1u << (nth % 8));
}

static void add_rule(const char *module, uint64_t off, int dev)
{
    if (g_nrules == g_rules_cap) {
        int nc = g_rules_cap ? g_rules_cap * 2 : 64;
        struct cs_rule *n = map_zeroed(nc * sizeof *n);
        if (!n) return;
        if (g_rules) {
            memcpy(n, g_rules, g_nrules * sizeof *n);
            munmap(g_rules, g_rules_cap * sizeof *n);
        }
        g_rules = n;
        g_rules_cap = nc;
    }
    snprintf(g_rules[g_nrules].module, sizeof g_rules[0].module, "%s", module);
    g_rules[g_nrules].offset = off;
    g_rules[g_nrules].device = dev;
    g_nrules++;
}

int gb_objtrack_load(const char *path)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    char *line = NULL;
    size_t cap = 0;
    while (getline(&line, &cap, f) > 0) {
        char c[2];
        uint64_t nth;
        char mod[128];
        if (sscanf(line, "%1[HD] #%" SCNu64, c, &nth) == 2) {
            set_bit(nth, c[0] == 'D');
        } else if (sscanf(line, "%1[HD] @%127[^+ \t\n]+0x%" SCNx64, c, mod, &nth) == 3) {
            add_rule(mod, nth, c[0] == 'D');
        }
    }
    free(line);
    int err = ferror(f);
    fclose(f);
    return err ? -1 : 0;
}

/* Return address -> "module+0xoffset".  dladdr() names the object that
 * contains the address; the main program may come back unnamed. */
void gb_callsite_name(void *ra, char *buf, size_t len)
{
    Dl_info info;
    if (ra && dladdr(ra, &info) && info.dli_fbase) {
        const char *name = info.dli_fname && *info.dli_fname ? info.dli_fname : NULL;
        if (!name) name = program_invocation_short_name;
        const char *base = strrchr(name, '/');
        base = base ? base + 1 : name;
        snprintf(buf, len, "%s+0x%" PRIx64, base, (uint64_t)((char *)ra - (char *)info.dli_fbase));
    } else {
        snprintf(buf, len, "?+0x%" PRIx64, (uint64_t)(uintptr_t)ra);
    }
}

/* Call-site decisions are cached per return address (lossy, lock-free). */
#define CS_CACHE 4096
static struct {
    void *ra;
    int decision; /* 1 host, 2 device */
} g_cs_cache[CS_CACHE];

static int callsite_decision(void *ra)
{
    if (!g_nrules) return 0;
    size_t h = ((uintptr_t)ra >> 2) % CS_CACHE;
    if (__atomic_load_n(&g_cs_cache[h].ra, __ATOMIC_ACQUIRE) == ra)
        return __atomic_load_n(&g_cs_cache[h].decision, __ATOMIC_RELAXED);
    char nm[192];
    gb_callsite_name(ra, nm, sizeof nm);
    int d = 1;
    for (int i = 0; i < g_nrules; ++i) {
        char want[192];
        snprintf(want, sizeof want, "%s+0x%" PRIx64, g_rules[i].module, g_rules[i].offset);
        if (!strcmp(want, nm)) {
            d = g_rules[i].device ? 2 : 1;
            break;
        }
    }
    __atomic_store_n(&g_cs_cache[h].decision, d, __ATOMIC_RELAXED);
    __atomic_store_n(&g_cs_cache[h].ra, ra, __ATOMIC_RELEASE);
    return d;
}

static uint64_t g_rand_state = 0x853c49e6748fea9bull;

bool gb_should_track(uint64_t nth, void *callsite, size_t size)
{
    if (size < gb_opt.min_track) return false;
    switch (gb_opt.heuristic) {
    case H_TRUE:
        return true;
    case H_RANDOM: {
        uint64_t x = __atomic_add_fetch(&g_rand_state, 0x9E3779B97F4A7C15ull, __ATOMIC_RELAXED);
        x = (x ^ (x >> 31)) * 0xBF58476D1CE4E5B9ull;
        return (x >> 63) != 0;
    }
    case H_SIZE:
        return size >= gb_opt.size_threshold;
    case H_OBJTRACK:
        if (nth < g_nbits && (g_bits[nth / 8] >> (nth % 8)) & 1) return true;
        return callsite_decision(callsite) == 2;
    default:
        return false;
    }
}

/* ---- trace records (T = tracked, U = untracked/freed, C = BLAS use) ---- */
static int g_trace_fd = -2;

static int trace_fd(void)
{
    if (g_trace_fd == -2) {
        g_trace_fd = -1;
        if (gb_opt.trace_file[0])
            g_trace_fd = open(gb_opt.trace_file, O_WRONLY | O_CREAT | O_APPEND, 0644);
    }
    return g_trace_fd;
}

void gb_trace_alloc(char kind, uint64_t nth, void *callsite, size_t size, void *p)
{
    if (!gb_opt.trace_file[0]) return;
    int fd = trace_fd();
    if (fd < 0) return;
    char cs[192];
    gb_callsite_name(callsite, cs, sizeof cs);
    gb_log(fd, "%c #%" PRIu64 " [%p] site=[%s] reqsize=[%zu]\n", kind,
           nth == UINT64_MAX ? (uint64_t)0 : nth, p, cs, size);
}

void gb_trace_call(const char *routine, const void *p)
{
    if (!gb_opt.trace_file[0]) return;
    int fd = trace_fd();
    if (fd < 0) return;
    struct gb_obj *o = gb_obj_find(p);
    if (!o) return;
    gb_log(fd, "C #%" PRIu64 " [%p] fun=[%s]\n", o->nth, p, routine);
}

/* ---- capture: allocations that were BLAS operands -> object-tracking file
 * Records come from freed objects (obj_free) and, at exit, from the live
 * ones.  The file holds one call-site rule per site (stable across runs and
 * library versions) and the ordinals (valid for an identical run), each with
 * the counts behind it as a comment.  Storage is mmap-backed: this runs
 * inside the allocator hooks. ---- */
struct use_rec {
    uint64_t nth;
    void *site;
    size_t size;
};
static struct use_rec *g_use = NULL;
static size_t g_use_n = 0, g_use_cap = 0;
static pthread_mutex_t g_use_mu = PTHREAD_MUTEX_INITIALIZER;

void gb_objtrack_note(uint64_t nth, void *callsite, size_t size)
{
    if (nth == UINT64_MAX) return; /* explicit gpublas_malloc: no ordinal, always tracked */
    pthread_mutex_lock(&g_use_mu);
    if (g_use_n == g_use_cap) {
        size_t nc = g_use_cap ? 2 * g_use_cap : 4096;
        struct use_rec *nt = map_zeroed(nc * sizeof *nt);
        if (!nt) {
            pthread_mutex_unlock(&g_use_mu);
            return;
        }
        if (g_use) {
            memcpy(nt, g_use, g_use_n * sizeof *nt);
            munmap(g_use, g_use_cap * sizeof *nt);
        }
        g_use = nt;
        g_use_cap = nc;
    }
    g_use[g_use_n++] = (struct use_rec){nth, callsite, size};
    pthread_mutex_unlock(&g_use_mu);
}

static int by_site_then_nth(const void *a, const void *b)
{
    const struct use_rec *x = a, *y = b;
    if (x->site != y->site) return (uintptr_t)x->site < (uintptr_t)y->site ? -1 : 1;
    return x->nth < y->nth ? -1 : x->nth > y->nth;
}

static int by_nth(const void *a, const void *b)
{
    const struct use_rec *x = a, *y = b;
    return x->nth < y->nth ? -1 : x->nth > y->nth;
}

void gb_objtrack_write(void)
{
    gb_obj_foreach_used(gb_objtrack_note);
    int fd = open(gb_opt.objtrack_out, O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (fd < 0) {
        gb_log(2, "libgpublas: cannot write object-tracking file '%s'\n", gb_opt.objtrack_out);
        return;
    }
    pthread_mutex_lock(&g_use_mu);
    gb_log(fd, "# object-tracking file written by libgpublas (objtrack_out) for %s:\n"
               "# allocations that were BLAS operands.  Load it with heuristic=objtrack:<file>.\n"
               "# Call-site rules (D @module+0xoffset) hold across runs; the ordinals (D #n)\n"
               "# only for an identical run.\n",
           program_invocation_short_name);
    qsort(g_use, g_use_n, sizeof *g_use, by_site_then_nth);
    for (size_t k = 0; k < g_use_n;) {
        size_t e = k;
        uint64_t bytes = 0;
        while (e < g_use_n && g_use[e].site == g_use[k].site) bytes += g_use[e++].size;
        char cs[192];
        gb_callsite_name(g_use[k].site, cs, sizeof cs);
        gb_log(fd, "D @%s # %zu allocation(s), %" PRIu64 " bytes\n", cs, e - k, bytes);
        k = e;
    }
    qsort(g_use, g_use_n, sizeof *g_use, by_nth);
    for (size_t k = 0; k < g_use_n; ++k) {
        if (k && g_use[k].nth == g_use[k - 1].nth) continue;
        char cs[192];
        gb_callsite_name(g_use[k].site, cs, sizeof cs);
        gb_log(fd, "D #%" PRIu64 " site=[%s] reqsize=[%zu]\n", g_use[k].nth, cs, g_use[k].size);
    }
    pthread_mutex_unlock(&g_use_mu);
    close(fd);
}
