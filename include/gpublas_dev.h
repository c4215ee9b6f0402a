/*
 * gpublas_dev.h -- the thin C ABI between the C host layer and the gfx950
 * HIP kernels.  It replaces the reference's vendor-BLAS shim: the cuBLAS /
 * clBLAS handle and calls behind runtime-blas.h:73-83 and the per-routine
 * call sites (blas_level3/gemm.cc:62-82, syrk.cc:57-75, trsm.cc:52-72,
 * symm.cc:57-78, and the unbuilt blas_level1/2 wrappers).
 *
 * Conventions: column-major, 0-based device pointers (names start with d_),
 * Fortran leading dimensions, trans/uplo/side/diag as the netlib characters
 * (case-insensitive).  `stream` is a hipStream_t or NULL for the calling
 * thread's own stream.  Arguments are assumed legal (the host layer checks
 * them netlib-style first).  Every function returns 0 on success or a
 * hipError_t code; launches are asynchronous on the stream.
 */
#ifndef GPUBLAS_DEV_H
#define GPUBLAS_DEV_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- device, streams, memory (replaces runtime.h:78-125) ---- */
int gpublas_dev_init(void);              /* 0 = a gfx950 device is ready */
int gpublas_dev_count(void);
const char *gpublas_dev_arch(void);      /* e.g. "gfx950:sramecc+:xnack-" */
size_t gpublas_dev_total_mem(void);
void *gpublas_dev_stream(void);          /* calling thread's stream */
int gpublas_dev_sync(void *stream);
/* Stream sequence numbers: mark() numbers the work enqueued so far on a
 * stream; wait(seq) returns once that work is complete (no-op if a sync has
 * already observed it).  Streams live for the whole process. */
uint64_t gpublas_dev_stream_mark(void *stream);
int gpublas_dev_stream_wait(void *stream, uint64_t seq);
int gpublas_dev_stream_done(void *stream, uint64_t seq);  /* non-blocking: 1 if complete */
/* The caller has seen work up to `seq` of the stream end (e.g. a reduction
 * published after it): later done/wait queries for it return at once. */
void gpublas_dev_stream_observed(void *stream, uint64_t seq);
/* In a child process after fork(): the GPU runtime is unusable there; mark
 * the device layer uninitialisable and forget the resident server without
 * any runtime call. */
void gpublas_dev_forked_child(void);
/* 1 if a synchronisation or observation has already seen `seq` complete (no query). */
int gpublas_dev_stream_known_done(void *stream, uint64_t seq);
/* Scalar results: the next sequence number of the calling thread's result
 * slot, and a poll for the reduction kernel's publication of that number. */
uint64_t gpublas_dev_scalar_next(void);
int gpublas_dev_scalar_wait(void *stream, uint64_t seq);
void *gpublas_dev_alloc(size_t bytes);   /* caching pool, 256-B aligned; NULL on OOM */
void gpublas_dev_release(void *d_ptr);   /* back to the pool; caller has synced users */
int gpublas_dev_trim(void);              /* return cached blocks to the driver */
size_t gpublas_dev_pool_bytes(void);     /* bytes held by the pool (in use + cached) */
int gpublas_dev_h2d(void *d_dst, const void *h_src, size_t bytes, void *stream);
int gpublas_dev_d2h(void *h_dst, const void *d_src, size_t bytes, void *stream);
int gpublas_dev_d2d(void *d_dst, const void *d_src, size_t bytes, void *stream);
/* 2-D copies of an m x n column-major block (element size es). */
int gpublas_dev_h2d_2d(void *d_dst, size_t ldd, const void *h_src, size_t lds, size_t m,
                       size_t n, size_t es, void *stream);
int gpublas_dev_d2h_2d(void *h_dst, size_t ldd, const void *d_src, size_t lds, size_t m,
                       size_t n, size_t es, void *stream);
/* Staged copies of an m x n column-major block between pageable host memory
 * and HBM: pieces of at most 4 MiB through a ring of pinned slots, host-side
 * packing on a small thread pool overlapped with the DMA (~57 GB/s link, vs
 * 17-36 GB/s for pageable copies; uploads under 1 MiB, where HIP's own
 * pageable copy is faster, go directly).  h2d returns with the DMAs enqueued
 * on the stream (the host block is no longer read); d2h returns with the data
 * in host memory.  Only the block's elements are written.  No counterpart in
 * the reference, whose per-call copies are plain cudaMemcpy
 * (blas_level3/gemm.cc:46-83). */
int gpublas_dev_h2d_staged(void *d_dst, size_t ldd, const void *h_src, size_t ldh, size_t m,
                           size_t n, size_t es, void *stream);
int gpublas_dev_d2h_staged(void *h_dst, size_t ldh, const void *d_src, size_t ldd, size_t m,
                           size_t n, size_t es, void *stream);
void *gpublas_dev_host_alloc(size_t bytes); /* pinned host memory */
void gpublas_dev_host_free(void *p);

/* Deterministic inputs: element i = 2*u-1, u = (splitmix64(seed,i)>>11)*2^-53. */
int gpublas_dev_fill_uniform(void *stream, uint64_t seed, int64_t n, double *d_out);
int gpublas_dev_fill_uniform_f32(void *stream, uint64_t seed, int64_t n, float *d_out);

/* Kernel timing with HIP events on the launch stream. */
void gpublas_dev_prof_enable(int on);
void gpublas_dev_prof_reset(void);
/* Resolve pending events (syncs their streams); returns the number of
 * launches of `kernel` recorded and writes the summed / min / max durations. */
int gpublas_dev_prof_query(const char *kernel, double *total_ms, double *min_ms, double *max_ms);
int gpublas_dev_prof_dump(char *buf, size_t len); /* "name count total_ms\n" lines */

/* ---- Level 3 ---- */
int gpublas_dev_dgemm(void *stream, char transa, char transb, int m, int n, int k, double alpha,
                      const double *d_A, int lda, const double *d_B, int ldb, double beta,
                      double *d_C, int ldc);
int gpublas_dev_sgemm(void *stream, char transa, char transb, int m, int n, int k, float alpha,
                      const float *d_A, int lda, const float *d_B, int ldb, float beta,
                      float *d_C, int ldc);
/* complex GEMM on interleaved (re, im) operands (replaces cublasZgemm /
 * cublasCgemm behind gemm.cc:181-217); ld counts complex elements; runs as one
 * real MFMA product of m x 2n x 2k plus pack and combine passes */
int gpublas_dev_zgemm(void *stream, char transa, char transb, int m, int n, int k,
                      double alpha_re, double alpha_im, const double *d_A, int lda,
                      const double *d_B, int ldb, double beta_re, double beta_im, double *d_C,
                      int ldc);
int gpublas_dev_cgemm(void *stream, char transa, char transb, int m, int n, int k,
                      float alpha_re, float alpha_im, const float *d_A, int lda, const float *d_B,
                      int ldb, float beta_re, float beta_im, float *d_C, int ldc);
/* the rest of the complex Level-3 (replaces the cuBLAS complex calls of
 * syrk.cc:182-215, herk.cc:131-162, syr2k.cc:182-215, her2k.cc:141-178,
 * symm.cc:158-207, hemm.cc:120-153, trsm.cc:175-215; ZTRMM / CTRMM follow netlib,
 * trmm.cc being unbuilt there): interleaved (re, im) operands, ld in complex
 * elements.  SYRK-type products run as two stacked real products on the
 * triangular tile grid; SYMM / HEMM and TRMM expand the triangle and use the
 * complex GEMM; TRSM is recursive with complex GEMM updates.  herm = 1 in
 * *symm selects HEMM. */
int gpublas_dev_zsyrk(void *stream, char uplo, char trans, int n, int k, double alpha_re,
                      double alpha_im, const double *d_A, int lda, double beta_re, double beta_im,
                      double *d_C, int ldc);
int gpublas_dev_csyrk(void *stream, char uplo, char trans, int n, int k, float alpha_re,
                      float alpha_im, const float *d_A, int lda, float beta_re, float beta_im,
                      float *d_C, int ldc);
int gpublas_dev_zherk(void *stream, char uplo, char trans, int n, int k, double alpha,
                      const double *d_A, int lda, double beta, double *d_C, int ldc);
int gpublas_dev_cherk(void *stream, char uplo, char trans, int n, int k, float alpha,
                      const float *d_A, int lda, float beta, float *d_C, int ldc);
int gpublas_dev_zsyr2k(void *stream, char uplo, char trans, int n, int k, double alpha_re,
                       double alpha_im, const double *d_A, int lda, const double *d_B, int ldb,
                       double beta_re, double beta_im, double *d_C, int ldc);
int gpublas_dev_csyr2k(void *stream, char uplo, char trans, int n, int k, float alpha_re,
                       float alpha_im, const float *d_A, int lda, const float *d_B, int ldb,
                       float beta_re, float beta_im, float *d_C, int ldc);
int gpublas_dev_zher2k(void *stream, char uplo, char trans, int n, int k, double alpha_re,
                       double alpha_im, const double *d_A, int lda, const double *d_B, int ldb,
                       double beta, double *d_C, int ldc);
int gpublas_dev_cher2k(void *stream, char uplo, char trans, int n, int k, float alpha_re,
                       float alpha_im, const float *d_A, int lda, const float *d_B, int ldb,
                       float beta, float *d_C, int ldc);
int gpublas_dev_zsymm(void *stream, int herm, char side, char uplo, int m, int n,
                      double alpha_re, double alpha_im, const double *d_A, int lda,
                      const double *d_B, int ldb, double beta_re, double beta_im, double *d_C,
                      int ldc);
int gpublas_dev_csymm(void *stream, int herm, char side, char uplo, int m, int n, float alpha_re,
                      float alpha_im, const float *d_A, int lda, const float *d_B, int ldb,
                      float beta_re, float beta_im, float *d_C, int ldc);
int gpublas_dev_ztrmm(void *stream, char side, char uplo, char transa, char diag, int m, int n,
                      double alpha_re, double alpha_im, const double *d_A, int lda, double *d_B,
                      int ldb);
int gpublas_dev_ctrmm(void *stream, char side, char uplo, char transa, char diag, int m, int n,
                      float alpha_re, float alpha_im, const float *d_A, int lda, float *d_B,
                      int ldb);
int gpublas_dev_ztrsm(void *stream, char side, char uplo, char transa, char diag, int m, int n,
                      double alpha_re, double alpha_im, const double *d_A, int lda, double *d_B,
                      int ldb);
int gpublas_dev_ctrsm(void *stream, char side, char uplo, char transa, char diag, int m, int n,
                      float alpha_re, float alpha_im, const float *d_A, int lda, float *d_B,
                      int ldb);
int gpublas_dev_dsyrk(void *stream, char uplo, char trans, int n, int k, double alpha,
                      const double *d_A, int lda, double beta, double *d_C, int ldc);
int gpublas_dev_ssyrk(void *stream, char uplo, char trans, int n, int k, float alpha,
                      const float *d_A, int lda, float beta, float *d_C, int ldc);
int gpublas_dev_dsymm(void *stream, char side, char uplo, int m, int n, double alpha,
                      const double *d_A, int lda, const double *d_B, int ldb, double beta,
                      double *d_C, int ldc);
int gpublas_dev_ssymm(void *stream, char side, char uplo, int m, int n, float alpha,
                      const float *d_A, int lda, const float *d_B, int ldb, float beta,
                      float *d_C, int ldc);
int gpublas_dev_dtrsm(void *stream, char side, char uplo, char transa, char diag, int m, int n,
                      double alpha, const double *d_A, int lda, double *d_B, int ldb);
int gpublas_dev_strsm(void *stream, char side, char uplo, char transa, char diag, int m, int n,
                      float alpha, const float *d_A, int lda, float *d_B, int ldb);
/* Rank-2k update and triangular multiply on the same tiles (netlib DSYR2K /
 * DTRMM semantics; the reference builds syr2k.cc:164-180 and leaves trmm.cc
 * disabled, meson.build:84-87).  TRMM uses two device workspaces of the
 * pool: the expanded triangle and the product before it is copied into B. */
int gpublas_dev_dsyr2k(void *stream, char uplo, char trans, int n, int k, double alpha,
                       const double *d_A, int lda, const double *d_B, int ldb, double beta,
                       double *d_C, int ldc);
int gpublas_dev_ssyr2k(void *stream, char uplo, char trans, int n, int k, float alpha,
                       const float *d_A, int lda, const float *d_B, int ldb, float beta,
                       float *d_C, int ldc);
int gpublas_dev_dtrmm(void *stream, char side, char uplo, char transa, char diag, int m, int n,
                      double alpha, const double *d_A, int lda, double *d_B, int ldb);
int gpublas_dev_strmm(void *stream, char side, char uplo, char transa, char diag, int m, int n,
                      float alpha, const float *d_A, int lda, float *d_B, int ldb);

/* Split-precision DGEMM / DSYRK (opt-in; no counterpart in the reference):
 * Ozaki integer-slice scheme, `slices` signed 8-bit slices per operand, slice
 * products p + q < slices on MFMA with FP32 accumulation per diagonal p + q
 * (exact for the leading diagonals), FP64 scaling and final sums.  Inputs
 * with Inf/NaN fall back to the FP64 kernels.  Synchronous.
 * Slice formats: BF16 (slices pre-scaled by 2^-8p; 1..8 slices) and FP16
 * (the integers themselves, each diagonal's weight applied in FP64; 1..6
 * slices); both give the same results at the same MFMA rate.  The entry
 * points without _fmt use BF16. */
#define GPUBLAS_SLICE_BF16 0
#define GPUBLAS_SLICE_FP16 1
/* INT8: signed 7-bit integer slices (1..8), products on
 * v_mfma_i32_32x32x32_i8 with exact int32 accumulation per diagonal p + q
 * (Ootomo, Ozaki & Yokota, IJHPCA 2024); s slices carry 7s bits. */
#define GPUBLAS_SLICE_INT8 2
int gpublas_dev_dgemm_split(void *stream, int slices, char transa, char transb, int m, int n,
                            int k, double alpha, const double *d_A, int lda, const double *d_B,
                            int ldb, double beta, double *d_C, int ldc);
int gpublas_dev_dsyrk_split(void *stream, int slices, char uplo, char trans, int n, int k,
                            double alpha, const double *d_A, int lda, double beta, double *d_C,
                            int ldc);
int gpublas_dev_dgemm_split_fmt(void *stream, int slices, int fmt, char transa, char transb, int m,
                                int n, int k, double alpha, const double *d_A, int lda,
                                const double *d_B, int ldb, double beta, double *d_C, int ldc);
int gpublas_dev_dsyrk_split_fmt(void *stream, int slices, int fmt, char uplo, char trans, int n,
                                int k, double alpha, const double *d_A, int lda, double beta,
                                double *d_C, int ldc);

/* ---- more Level 1 (netlib semantics: DSCAL and DASUM do nothing / return 0
 * for incx <= 0; negative increments walk the vectors backwards) ---- */
int gpublas_dev_dscal(void *stream, int n, double alpha, double *d_x, int incx);
int gpublas_dev_dcopy(void *stream, int n, const double *d_x, int incx, double *d_y, int incy);
int gpublas_dev_dswap(void *stream, int n, double *d_x, int incx, double *d_y, int incy);
int gpublas_dev_drot(void *stream, int n, double *d_x, int incx, double *d_y, int incy, double c,
                     double s);
int gpublas_dev_dasum(void *stream, int n, const double *d_x, int incx, double *h_result);

/* ---- Level 2 ---- */
int gpublas_dev_dgemv(void *stream, char trans, int m, int n, double alpha, const double *d_A,
                      int lda, const double *d_x, int incx, double beta, double *d_y, int incy);
/* Allocation-free DGEMV.  No counterpart in the reference: it is for callers
 * that cannot allocate -- code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  Same arguments, kernels and results (bit
 * for bit) as gpublas_dev_dgemv, but it allocates nothing: the column-panel
 * partials of trans 'N' live in a fixed 4 MiB device array per GPU (2^19
 * doubles, enough for every shape; held until the process ends), whose
 * address gpublas_dev_init() takes once per device.  Differences: `stream`
 * is used as given (NULL = the device's null stream, not the calling
 * thread's stream), no kernel timing is recorded, and it returns
 * hipErrorNotInitialized if gpublas_dev_init() has not run (call it once
 * first) and hipErrorInvalidValue if a shape ever needed more partials than
 * the array holds.  Every call on one GPU shares that array: calls that
 * overlap in time on one GPU (two streams, two threads) race and must be
 * ordered by the caller. */
int gpublas_dev_dgemv_noalloc(void *stream, char trans, int m, int n, double alpha,
                              const double *d_A, int lda, const double *d_x, int incx, double beta,
                              double *d_y, int incy);
int gpublas_dev_dger(void *stream, int m, int n, double alpha, const double *d_x, int incx,
                     const double *d_y, int incy, double *d_A, int lda);
int gpublas_dev_dtrsv(void *stream, char uplo, char trans, char diag, int n, const double *d_A,
                      int lda, double *d_x, int incx);
/* DSYMV / DSYR on the uplo triangle (replace cublasDsymv, blas_level2/symv.cc:57-64,
 * and cublasDsyr, syr.cc:50-55). */
int gpublas_dev_dsymv(void *stream, char uplo, int n, double alpha, const double *d_A, int lda,
                      const double *d_x, int incx, double beta, double *d_y, int incy);
int gpublas_dev_dsyr(void *stream, char uplo, int n, double alpha, const double *d_x, int incx,
                     double *d_A, int lda);
/* DSYR2 (replaces cublasDsyr2, blas_level2/syr2.cc) and the packed rank
 * updates DSPR / DSPR2 (cublasDspr / cublasDspr2, spr.cc, spr2.cc) on the
 * netlib packed layout. */
int gpublas_dev_dsyr2(void *stream, char uplo, int n, double alpha, const double *d_x, int incx,
                      const double *d_y, int incy, double *d_A, int lda);
int gpublas_dev_dspr(void *stream, char uplo, int n, double alpha, const double *d_x, int incx,
                     double *d_AP);
int gpublas_dev_dspr2(void *stream, char uplo, int n, double alpha, const double *d_x, int incx,
                      const double *d_y, int incy, double *d_AP);
/* DTRMV, DTPMV, DTPSV, DSPMV (replace cublasDtrmv / Dtpmv / Dtpsv / Dspmv,
 * blas_level2/trmv.cc, tpmv.cc, tpsv.cc, pmv.cc); packed = netlib packing. */
int gpublas_dev_dtrmv(void *stream, char uplo, char trans, char diag, int n, const double *d_A,
                      int lda, double *d_x, int incx);
int gpublas_dev_dtpmv(void *stream, char uplo, char trans, char diag, int n, const double *d_AP,
                      double *d_x, int incx);
int gpublas_dev_dtpsv(void *stream, char uplo, char trans, char diag, int n, const double *d_AP,
                      double *d_x, int incx);
int gpublas_dev_dspmv(void *stream, char uplo, int n, double alpha, const double *d_AP,
                      const double *d_x, int incx, double beta, double *d_y, int incy);
/* DGBMV, DSBMV, DTBMV, DTBSV (replace cublasDgbmv / Dsbmv / Dtbmv / Dtbsv,
 * blas_level2/gbmv.cc, sbmv.cc, tbmv.cc, tbsv.cc); netlib band storage.
 * DTBSV supports k <= 8191 (hipErrorInvalidValue beyond). */
int gpublas_dev_dgbmv(void *stream, char trans, int m, int n, int kl, int ku, double alpha,
                      const double *d_A, int lda, const double *d_x, int incx, double beta,
                      double *d_y, int incy);
int gpublas_dev_dsbmv(void *stream, char uplo, int n, int k, double alpha, const double *d_A,
                      int lda, const double *d_x, int incx, double beta, double *d_y, int incy);
int gpublas_dev_dtbmv(void *stream, char uplo, char trans, char diag, int n, int k,
                      const double *d_A, int lda, double *d_x, int incx);
int gpublas_dev_dtbsv(void *stream, char uplo, char trans, char diag, int n, int k,
                      const double *d_A, int lda, double *d_x, int incx);
/* DTBSV plan switch (no reference counterpart): on = 1 (the default) lets
 * k <= 63, n >= 2048 solves run chunked (per-chunk responses, a propagation
 * of the incoming unknowns, per-chunk re-solves; declined on the device, in
 * favour of the one-wave recurrence, when a response leaves [-1, 1] or a
 * value is not finite); on = 0 keeps every solve on the one-wave
 * recurrence; on < 0 only queries.  Returns the previous setting. */
int gpublas_dev_tbsv_chunked(int on);

/* ---- Level 1 (scalar results land in *h_result, pinned or device-visible host
 * memory from gpublas_dev_host_alloc, valid after the stream syncs) ---- */
int gpublas_dev_daxpy(void *stream, int n, double alpha, const double *d_x, int incx,
                      double *d_y, int incy);
int gpublas_dev_ddot(void *stream, int n, const double *d_x, int incx, const double *d_y,
                     int incy, double *h_result);
int gpublas_dev_dnrm2(void *stream, int n, const double *d_x, int incx, double *h_result);
int gpublas_dev_idamax(void *stream, int n, const double *d_x, int incx, int *h_result);
/* IDAMIN (replaces cublasIdamin, blas_level1/amin.cc:40) and DROTM (replaces
 * cublasDrotm, rotm.cc:51; h_param = flag, h11, h21, h12, h22 on the host). */
int gpublas_dev_idamin(void *stream, int n, const double *d_x, int incx, int *h_result);
int gpublas_dev_drotm(void *stream, int n, double *d_x, int incx, double *d_y, int incy,
                      const double *h_param);
/* Pinned result slot of the calling thread (256 bytes). */
void *gpublas_dev_result_slot(void);
/* Resident reduction server (opt-in; no counterpart in the reference, which
 * launches one vendor call per intercepted call, runtime-blas.c:14-20): a
 * kernel that stays on the device (two 1024-thread workgroups) and serves
 * short reductions (n <= 65536) from a ring of request slots in pinned host
 * memory, so a scalar-returning call pays no kernel launch.  reduce() runs
 * op 0 DDOT, 1 DNRM2, 2 IDAMAX, 3 IDAMIN, 4 DASUM (1-based index in *ires
 * for 2 and 3, netlib rules) and returns once the result is in *dres /
 * *ires.  The operands must not be pending writes: with operands_ready = 0
 * the call is served only if `stream` has no work queued; operands_ready = 1
 * asserts that the kernels that wrote them have ended.  It returns -1
 * (nothing done) when the server is not running, another request is in
 * flight or the call is not one it serves, so the caller can take the
 * launched path.  Results are bit-identical to the
 * launched kernels'.  The server exits at process exit or on stop(). */
int gpublas_dev_resident_start(void);
void gpublas_dev_resident_stop(void);
int gpublas_dev_resident_active(void);
int gpublas_dev_resident_reduce(void *stream, int op, int n, const double *d_x, int incx,
                                const double *d_y, int incy, int operands_ready, double *dres,
                                int *ires);

/* ---- Batch slots of the serving loop (no counterpart in the reference,
 * which issues one cuBLAS call per intercepted call: blas_level3/gemm.cc:62-82,
 * trsm.cc:52-72).  A slot collects up to 24 GEMMs per tile shape and
 * precision, 32 DGEMVs and 48 DTRSMs (m <= 1024), of any sizes, and runs
 * them with one grouped persistent launch per kind -- one
 * replay of a captured hipGraph (use_graph != 0) or the same kernels
 * launched eagerly.  Operands are device pointers.  add_*: 0 = added,
 * 1 = slot full (launch it and use another), -1 = not supported in a batch
 * (GEMM with alpha == 0 or k == 0, right-side TRSM, zero strides): the
 * caller uses the per-call entry points.  Results do not depend on what
 * else is in the batch.  A slot is used by one host thread at a time; clear
 * it only after done() reports completion. ---- */
typedef struct gpublas_batch gpublas_batch;
gpublas_batch *gpublas_dev_batch_create(int use_graph);
void gpublas_dev_batch_destroy(gpublas_batch *b);
void gpublas_dev_batch_clear(gpublas_batch *b);
int gpublas_dev_batch_count(const gpublas_batch *b);
void *gpublas_dev_batch_stream(gpublas_batch *b);
/* -1 = tile shape from the tuned table (default); 0/1/2 = 32/64/128 square tiles */
void gpublas_dev_batch_force_cfg(gpublas_batch *b, int cfg);
int gpublas_dev_batch_dgemm(gpublas_batch *b, char transa, char transb, int m, int n, int k,
                            double alpha, const double *d_A, int lda, const double *d_B, int ldb,
                            double beta, double *d_C, int ldc);
int gpublas_dev_batch_sgemm(gpublas_batch *b, char transa, char transb, int m, int n, int k,
                            float alpha, const float *d_A, int lda, const float *d_B, int ldb,
                            float beta, float *d_C, int ldc);
int gpublas_dev_batch_dgemv(gpublas_batch *b, char trans, int m, int n, double alpha,
                            const double *d_A, int lda, const double *d_x, int incx, double beta,
                            double *d_y, int incy);
int gpublas_dev_batch_dtrsm(gpublas_batch *b, char side, char uplo, char transa, char diag, int m,
                            int n, double alpha, const double *d_A, int lda, double *d_B, int ldb);
int gpublas_dev_batch_launch(gpublas_batch *b);   /* asynchronous */
int gpublas_dev_batch_done(gpublas_batch *b);     /* 1 once the launched batch has completed */
/* 1 once the batch's problems of one kind (0 DGEMM, 1 SGEMM, 2 DGEMV,
 * 3 DTRSM) have completed; eager launches run the kinds in increasing
 * estimated cost, so cheap kinds finish first (graph replays: whole batch) */
int gpublas_dev_batch_done_kind(gpublas_batch *b, int kind);
int gpublas_dev_batch_wait(gpublas_batch *b);
/* GEMM tile-shape table of the batch kernels: load from `path` (unless
 * force), else measure every tile shape per (sqrt(mn), k) size bucket and
 * write the table to `path` (NULL: keep in memory only). */
int gpublas_dev_tune_gemm(const char *path, int force);
int gpublas_dev_tune_choice(int is_single, int m, int n, int k);
void gpublas_dev_tune_reset(void);

#ifdef __cplusplus
}
#endif
#endif
