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
#include <signal.h>
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
    int split_fmt;           /* GPUBLAS_SLICE_BF16 (default) or GPUBLAS_SLICE_FP16 */
    int serve;               /* route calls on tracked operands through the serving loop */
    int resident;            /* short reductions on tracked operands through the resident server */
    size_t serve_budget;     /* HBM admission budget of the serving loop (bytes) */
    int serve_slots;         /* batches in flight */
    int serve_graph;         /* replay captured graphs (else launch eagerly) */
    int serve_batch;         /* most requests per batch */
    char tune_file[256];     /* GEMM tile-shape table of the batch kernels */
    int verbose;
    char stats_file[256];    /* "" = no statistics file at exit */
    char trace_file[256];    /* "" = no allocation / call trace */
    char objtrack_out[256];  /* "" = no object-tracking file written at exit */
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

/* ---- host or GPU for calls with no resident operand (route.c) ---- */
#define GB_CUTOFF_AUTO (-1) /* cutoff=auto: the measured crossover table */
enum gb_route_kind {
    GB_RK_DGEMM, GB_RK_SGEMM, GB_RK_DSYRK, GB_RK_SSYRK, GB_RK_DSYR2K, GB_RK_SSYR2K,
    GB_RK_DSYMM, GB_RK_SSYMM, GB_RK_DTRSM, GB_RK_STRSM, GB_RK_DTRMM, GB_RK_STRMM,
    GB_RK_DGEMV, GB_RK_DGER, GB_RK_DTRSV, GB_RK_DSYMV, GB_RK_DSYR,
    GB_RK_STREAM, /* daxpy, dscal, dcopy, dswap, drot */
    GB_RK_REDUCE, /* ddot, dnrm2, dasum, idamax */
    GB_RK_COUNT
};
/* 1 = run the call on the host BLAS.  below_fixed is the fixed rule's verdict
 * (cutoff=<n>); flops and bytes feed the crossover table (cutoff=auto); a
 * tracked p, q or r (NULL allowed) keeps the call on the GPU. */
int gb_host_route(int kind, int below_fixed, double flops, double bytes, const void *p,
                  const void *q, const void *r);
int gb_xover_load(const char *path);
extern GB_TLS int gb_route_gpu; /* set by the crossover measurement: force the GPU */

/* ---- device availability: aborts with a message if no gfx950 device ---- */
void gb_require_device(void);

/* ---- tracked objects (tracker.c) ---- */
/* ---- object-tracking capture (objtrack.c): allocations used as BLAS
 * operands, written at exit as an object-tracking file ---- */
void gb_objtrack_note(uint64_t nth, void *callsite, size_t size);
void gb_objtrack_write(void);
void gb_obj_foreach_used(void (*fn)(uint64_t nth, void *callsite, size_t size));

enum gb_state { ST_HOST = 0, ST_SHARED = 1, ST_DEVICE = 2 };
#define GB_RD_SLOTS 32 /* reader streams remembered per object */
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
    void *wr_stream;     /* stream of the last kernel that wrote the mirror */
    uint64_t wr_seq;     /* its sequence number on that stream */
    void *rd_stream[GB_RD_SLOTS];  /* kernels reading the mirror since, one per stream */
    uint64_t rd_seq[GB_RD_SLOTS];
    int nrd;
    int refs;            /* calls between acquire and finish, served requests in flight */
    uint64_t last_use;   /* acquire tick, for least-recently-used eviction */
    int detached;        /* in a fork child: plain private memory, no longer tracked */
    pthread_mutex_t mu;
};
void *gb_obj_create(size_t size, uint64_t nth, void *callsite, int zero);
int gb_obj_destroy(void *p);                 /* 1 if p was a tracked object */
void *gb_obj_realloc(void *p, size_t size);  /* p tracked */
struct gb_obj *gb_obj_find(const void *p);   /* tracked object containing p, or NULL */
int gb_obj_owns(const void *p);              /* p starts an allocation of this library */
size_t gb_obj_usable_size(const void *p);    /* 0 if untracked */
int gb_tracked_count(void);
void gb_install_fault_handler(void);
/* Make [p, p+n) usable by a system call or a library routine that does not
 * fault into the SIGSEGV handler: every tracked object the range overlaps is
 * brought to the host (its device copy downloaded if newer) and its view
 * opened for reading, or for writing (demoting it to HOST) when `write`. */
void gb_host_access(const void *p, size_t n, int write);
/* The application's sigaction() / signal() on SIGSEGV once the library's
 * handler is installed: `act` becomes the handler the library chains to
 * for faults that are not its own, `old` receives the previous one.
 * Returns 0 if the library's handler is not installed (the caller then
 * performs the call itself). */
int gb_segv_app_action(const struct sigaction *act, struct sigaction *old);
/* Reinstall the library's SIGSEGV handler if something replaced it without
 * going through sigaction()/signal() (checked at most once a millisecond). */
void gb_check_fault_handler(void);
/* The libc sigaction, bypassing the interposer. */
int gb_real_sigaction(int sig, const struct sigaction *act, struct sigaction *old);
/* A libc function by name, bypassing any interposer (NULL if absent). */
void *gb_libc_sym(const char *name);
/* Set in a child process after fork(): every BLAS call runs on the host. */
extern int gb_forked_child;
/* Make the device copy valid (upload if the host copy is newer); returns 1 if
 * it uploaded.  Takes a reference (released by gb_finish / gb_obj_unref).
 * Orders `stream` after conflicting kernels on other streams: the last
 * writer, and for a write (or an upload) every recorded reader too. */
int gb_obj_to_device(struct gb_obj *o, void *stream, int write);
void gb_obj_unref(struct gb_obj *o);
/* Wait until no kernel uses the object's device mirror any more. */
void gb_obj_quiesce(struct gb_obj *o);
/* Wait for the kernels a use of the object must follow (host-side). */
void gb_obj_wait_hazards(struct gb_obj *o, int write);
/* HBM held by device mirrors of tracked objects. */
extern size_t gb_resident_bytes;
/* Drop the device mirrors of idle objects whose host copy is current (least
 * recently used first) until `want` bytes are freed; objects in keep[] are
 * spared.  Returns the bytes freed. */
size_t gb_evict_idle(size_t want, struct gb_obj *const *keep, int nkeep);
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
/* Elements of a packed triangle of order n; n(n+1)/2 passes INT_MAX from
 * n = 65536, so it is computed in size_t. */
static inline size_t gb_packed_len(int n)
{
    return n > 0 ? (size_t)n * ((size_t)n + 1) / 2 : 0;
}
/* Packed triangle of order n: one contiguous span of gb_packed_len(n) elements. */
void gb_acquire_packed(gb_operand *op, const void *host, int n, size_t es, int mode, void *stream);
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
/* Every operand is a tracked object whose last writing kernel has completed. */
bool gb_writers_done(const gb_operand *ops, int nops);
/* End of a reduction served by the resident server (result already in hand). */
void gb_finish_served(void *stream, gb_operand *ops, int nops);

/* ---- serving loop (serve.c) ---- */
/* Each returns 1 if the call was served (all operands tracked, shape
 * supported), 0 if the caller must take the per-call path. */
int gb_serve_dgemm(char ta, char tb, int m, int n, int k, double alpha, const double *a, int lda,
                   const double *b, int ldb, double beta, double *c, int ldc);
int gb_serve_sgemm(char ta, char tb, int m, int n, int k, float alpha, const float *a, int lda,
                   const float *b, int ldb, float beta, float *c, int ldc);
int gb_serve_dgemv(char trans, int m, int n, double alpha, const double *a, int lda,
                   const double *x, int incx, double beta, double *y, int incy);
int gb_serve_dtrsm(char side, char uplo, char transa, char diag, int m, int n, double alpha,
                   const double *a, int lda, double *b, int ldb);
void gb_serve_stop(void);
bool gb_ptr_tracked(const void *p);

void gb_check(int err, const char *what);

#endif
