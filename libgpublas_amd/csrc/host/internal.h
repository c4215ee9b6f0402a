/*
 * internal.h -- shared state of the C host layer (not installed).
 */
#ifndef GPUBLAS_INTERNAL_H
#define GPUBLAS_INTERNAL_H
#define _GNU_SOURCE
#include <stddef.h>
#include <stdint.h>
#include <stdbool.h>
#include <pthread.h>
#include "gpublas.h"
#include "gpublas_dev.h"

#define GB_TLS __thread __attribute__((tls_model("initial-exec")))

/* ---- logging without malloc (raw write(2); reference common.h:15-28) ---- */
void gb_log(int fd, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void gb_fatal(const char *fmt, ...) __attribute__((format(printf, 1, 2), noreturn));

/* ---- re-entrancy guard: set while any library code runs on this thread,
 * so allocations made by the library or the HIP runtime on its behalf are
 * never tracked (reference obj_tracker.c:343-350 inside_internal). ---- */
extern GB_TLS int gb_guard;
void gb_init(void);
static inline void gb_enter(void) { gb_guard++; }
static inline void gb_leave(void) { gb_guard--; }

/* ---- options (reference blas2cuda.c:46-124) ---- */
enum gb_heuristic { H_FALSE = 0, H_TRUE, H_RANDOM, H_SIZE, H_OBJTRACK };
struct gb_options {
    enum gb_heuristic heuristic;
    size_t size_threshold;   /* H_SIZE: track allocations >= this */
    size_t min_track;        /* never track allocations below this */
    int cutoff;              /* gemm fall-through cutoff on leading dims */
    int profile;             /* kernel event timing */
    int split;               /* >0: DGEMM/DSYRK as split-precision with this many slices */
    int verbose;
    char stats_file[256];    /* "" = no statistics file at exit */
    char trace_file[256];    /* "" = no allocation / call trace */
};
extern struct gb_options gb_opt;
int gb_parse_options(const char *s);

/* ---- statistics ---- */
extern gpublas_stats_t gb_stats;
#define GB_COUNT(field, v) __atomic_fetch_add(&gb_stats.field, (uint64_t)(v), __ATOMIC_RELAXED)

/* ---- xerbla / lsame (runtime-blas.c:34-58) ---- */
void gb_xerbla(const char *name6, int info);
int gb_lsame(char a, char b);

/* ---- host BLAS fall-through (runtime-blas.c:60-69) ---- */
void *gb_host_fn(const char *f77name);

/* ---- device availability: aborts with a message if no gfx950 device ---- */
void gb_require_device(void);

/* ---- tracked objects (tracker.c) ---- */
enum gb_state { ST_HOST = 0, ST_SHARED = 1, ST_DEVICE = 2 };
struct gb_obj {
    char *host;          /* application view (page aligned) */
    char *alias;         /* library view of the same pages, always RW */
    char *dev;           /* device mirror, NULL until first device use */
    size_t size;         /* requested bytes */
    size_t map_size;     /* page-rounded mapping */
    uint64_t nth;        /* allocation ordinal, UINT64_MAX if explicit */
    uint64_t uid;
    void *callsite;
    int state;
    int used;            /* seen as a BLAS operand */
    void *use_stream;    /* stream of the last kernel that used the mirror */
    uint64_t use_seq;    /* its sequence number on that stream */
    pthread_mutex_t mu;
};
void *gb_obj_create(size_t size, uint64_t nth, void *callsite, int zero);
int gb_obj_destroy(void *p);                 /* 1 if p was a tracked object */
void *gb_obj_realloc(void *p, size_t size);  /* p tracked */
struct gb_obj *gb_obj_find(const void *p);   /* object containing p, or NULL */
size_t gb_obj_usable_size(const void *p);    /* 0 if untracked */
int gb_tracked_count(void);
void gb_install_fault_handler(void);
/* Make the device copy valid (upload if the host copy is newer); returns 1 if
 * it uploaded.  Orders `stream` after the object's last use elsewhere. */
int gb_obj_to_device(struct gb_obj *o, void *stream);
/* Wait until no kernel uses the object's device mirror any more. */
void gb_obj_quiesce(struct gb_obj *o);
/* Mark the device copy as the only valid one (after a kernel wrote it). */
void gb_obj_device_written(struct gb_obj *o);

/* ---- allocation decisions (objtrack.c; reference lib/oracle.c) ---- */
int gb_objtrack_load(const char *path);
bool gb_should_track(uint64_t nth, void *callsite, size_t size);
void gb_callsite_name(void *ra, char *buf, size_t len);
extern uint64_t gb_nallocs;
extern int gb_tracking;   /* interposer active */
void gb_trace_alloc(char kind, uint64_t nth, void *callsite, size_t size, void *p);
void gb_trace_call(const char *routine, const void *p);

/* ---- operand residency (the reference's gpuptr, runtime-mem.hpp:19-176) ---- */
enum { GB_IN = 1, GB_OUT = 2, GB_INOUT = 3 };
typedef struct {
    const char *host;
    char *dev;
    struct gb_obj *obj;  /* tracked: used in place */
    char *tmp;           /* untracked: pool block */
    size_t ld, rows, cols, es;
    int mode;
    int uploaded;        /* this call copied host data in (object or staging) */
} gb_operand;
/* Matrix rows x cols with leading dimension ld (elements of es bytes). */
void gb_acquire_matrix(gb_operand *op, const void *host, size_t ld, size_t rows, size_t cols,
                       size_t es, int mode, void *stream);
/* Vector of n elements with stride inc (span 1+(n-1)|inc|). */
void gb_acquire_vector(gb_operand *op, const void *host, int n, int inc, size_t es, int mode,
                       void *stream);
/* Copy outputs back (untracked) or mark them device-resident (tracked).
 * Call after the kernel launch. */
void gb_release(gb_operand *op, void *stream);
/* End of a void BLAS call.  When every operand is a tracked object that was
 * already resident, the call returns without synchronising: nothing the host
 * can see is pending (a host touch of a written object faults and waits for
 * the kernel first).  Otherwise it synchronises and frees staging buffers. */
void gb_finish(void *stream, gb_operand *ops, int nops);
/* End of a scalar-returning call: wait for the kernel's publication of `seq`
 * in the result slot, then release staging buffers. */
void gb_finish_scalar(void *stream, gb_operand *ops, int nops, uint64_t seq);
bool gb_any_tracked(const gb_operand *ops, int nops);
bool gb_ptr_tracked(const void *p);

void gb_check(int err, const char *what);

#endif
