// blas2.hip -- Level-2 kernels on gfx950: DGEMV (N and T), DGER, DTRSV.
//
// Replaces the vendor calls of the reference's unbuilt CBLAS wrappers:
// cublasDgemv (blas_level2/gemv.cc:97), cublasDger (ger.cc:92),
// cublasDtrsv (trsv.cc:86).
//
// DGEMV 'N' is the HBM stream of the Level-1/2 chain: y = alpha*A*x + beta*y
// with A column-major.  Two launches:
//  * gemv_n_stream: a workgroup owns a 512-row x CB-column panel of A; wave
//    w reads rows [128w, 128w+128) of every column of the panel, 16 B per
//    lane, so each load instruction is one contiguous 1 KB burst and the 4
//    waves together stream 4 KB of each column.  Eight columns' loads are in
//    flight per wave (non-temporal: A is read once), two workgroups per CU.
//    Each workgroup writes its 512 partial sums to a per-panel slot;
//  * gemv_n_fold: sums the column panels of each row in panel order and
//    applies alpha / beta.
// A single-launch variant that folds in the last workgroup of each row block
// needs a device-scope release per workgroup, which writes back the XCD's L2
// and measured slower at this grid size, so the fold is a second kernel.  The
// summation order depends only on (m, n): results are reproducible.
// DGEMV 'T' runs one wave per column of A (a contiguous dot product), eight
// 16-B loads in flight per lane and a wave64 shuffle tree.
#include "devcommon.hpp"
#include "gpublas_dev.h"

int gpublas_dev_dtrsm(void *stream, char side, char uplo, char transa, char diag, int m, int n,
                      double alpha, const double *d_A, int lda, double *d_B, int ldb);

namespace {

typedef double d2 __attribute__((ext_vector_type(2)));
constexpr int GN_ROWS = 512;  // rows per workgroup panel (256 threads x 2)
constexpr int GN_U = 8;       // columns in flight per wave

__global__ __launch_bounds__(256) void gemv_n_stream(int m, int n, double alpha, const double *A,
                                                     long lda, const double *x, long incx, long x0,
                                                     double beta, double *y, long incy, long y0,
                                                     int nrb, int cbw, int direct, int vec,
                                                     double *partials)
{
    const int rb = blockIdx.x % nrb, cb = blockIdx.x / nrb;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long row = (long)rb * GN_ROWS + wave * 128 + 2 * lane;
    const int c0 = cb * cbw;
    const int c1 = min(n, c0 + cbw);
    double s0 = 0, s1 = 0;
    if (vec && row + 1 < m) {
        const double *a = A + row;
        int c = c0;
        for (; c + GN_U <= c1; c += GN_U) {
            d2 v[GN_U];
#pragma unroll
            for (int u = 0; u < GN_U; ++u)
                v[u] = __builtin_nontemporal_load((const d2 *)(a + (long)(c + u) * lda));
#pragma unroll
            for (int u = 0; u < GN_U; ++u) {
                const double xv = x[x0 + (long)(c + u) * incx];
                s0 = fma(v[u].x, xv, s0);
                s1 = fma(v[u].y, xv, s1);
            }
        }
        for (; c < c1; ++c) {
            const d2 v = *(const d2 *)(a + (long)c * lda);
            const double xv = x[x0 + (long)c * incx];
            s0 = fma(v.x, xv, s0);
            s1 = fma(v.y, xv, s1);
        }
    } else {
        for (int c = c0; c < c1; ++c) {
            const double xv = x[x0 + (long)c * incx];
            if (row < m) s0 = fma(A[row + (long)c * lda], xv, s0);
            if (row + 1 < m) s1 = fma(A[row + 1 + (long)c * lda], xv, s1);
        }
    }
    if (direct) {
        if (row < m) {
            double *yp = y + y0 + row * incy;
            *yp = beta == 0.0 ? alpha * s0 : fma(beta, *yp, alpha * s0);
        }
        if (row + 1 < m) {
            double *yp = y + y0 + (row + 1) * incy;
            *yp = beta == 0.0 ? alpha * s1 : fma(beta, *yp, alpha * s1);
        }
        return;
    }
    double *pp = partials + (long)cb * m;
    if (row < m) pp[row] = s0;
    if (row + 1 < m) pp[row + 1] = s1;
}

__global__ __launch_bounds__(256) void gemv_n_fold(int m, int ncb, const double *partials,
                                                   double alpha, double beta, double *y, long incy,
                                                   long y0)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    double s = 0;
    int b = 0;
    for (; b + 4 <= ncb; b += 4) {
        const double p0 = partials[(long)b * m + i], p1 = partials[(long)(b + 1) * m + i];
        const double p2 = partials[(long)(b + 2) * m + i], p3 = partials[(long)(b + 3) * m + i];
        s += p0;
        s += p1;
        s += p2;
        s += p3;
    }
    for (; b < ncb; ++b) s += partials[(long)b * m + i];
    double *yp = y + y0 + i * incy;
    *yp = beta == 0.0 ? alpha * s : fma(beta, *yp, alpha * s);
}

__global__ __launch_bounds__(256) void gemv_t_kernel(int m, int n, double alpha, const double *A,
                                                     long lda, const double *x, long incx, long x0,
                                                     double beta, double *y, long incy, long y0,
                                                     int vec)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (long j = (long)blockIdx.x * 4 + wave; j < n; j += (long)gridDim.x * 4) {
        const double *a = A + j * lda;
        double s = 0;
        if (vec && incx == 1 && x0 == 0) {
            int i = 2 * lane;
            for (; i + 128 * 7 + 1 < m; i += 128 * 8) {
                d2 av[8], xv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    av[u] = __builtin_nontemporal_load((const d2 *)(a + i + 128 * u));
                    xv[u] = *(const d2 *)(x + i + 128 * u);
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    s = fma(av[u].x, xv[u].x, s);
                    s = fma(av[u].y, xv[u].y, s);
                }
            }
            for (; i + 1 < m; i += 128) {
                const d2 av = *(const d2 *)(a + i);
                const d2 xv = *(const d2 *)(x + i);
                s = fma(av.x, xv.x, s);
                s = fma(av.y, xv.y, s);
            }
            if (i < m) s = fma(a[i], x[i], s);
        } else {
            for (int i = lane; i < m; i += 64) s = fma(a[i], x[x0 + (long)i * incx], s);
        }
        s = wave_sum(s);
        if (lane == 0) {
            double *yp = y + y0 + j * incy;
            *yp = beta == 0.0 ? alpha * s : fma(beta, *yp, alpha * s);
        }
    }
}

__global__ void scal_vec_kernel(int n, double beta, double *y, long incy, long y0)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (long)gridDim.x * blockDim.x) {
        double *p = y + y0 + i * incy;
        *p = beta == 0.0 ? 0.0 : beta * *p;
    }
}

__global__ __launch_bounds__(256) void ger_kernel(int m, int n, double alpha, const double *x,
                                                  long incx, long x0, const double *y, long incy,
                                                  long y0, double *A, long lda)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    const double xi = x[x0 + i * incx];
    const int j0 = blockIdx.y * 16, j1 = min(n, j0 + 16);
    for (int j = j0; j < j1; ++j) {
        const double yj = y[y0 + (long)j * incy];
        if (yj != 0.0) {
            double *p = A + i + (long)j * lda;
            *p = fma(xi, alpha * yj, *p);
        }
    }
}

__global__ void gather_kernel(int n, const double *x, long incx, long x0, double *t)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (long)gridDim.x * blockDim.x) t[i] = x[x0 + i * incx];
}

__global__ void scatter_kernel(int n, const double *t, double *x, long incx, long x0)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (long)gridDim.x * blockDim.x) x[x0 + i * incx] = t[i];
}

inline long start_of(int n, int inc) { return inc < 0 ? -(long)(n - 1) * inc : 0; }
inline bool isN(char c) { return c == 'N' || c == 'n'; }

}  // namespace

extern "C" {

int gpublas_dev_dgemv(void *stream, char trans, int m, int n, double alpha, const double *d_A,
                      int lda, const double *d_x, int incx, double beta, double *d_y, int incy)
{
    hipStream_t s = pick_stream(stream);
    if (m <= 0 || n <= 0) return 0;
    const bool nt = isN(trans);
    const int lenx = nt ? n : m, leny = nt ? m : n;
    const long x0 = start_of(lenx, incx), y0 = start_of(leny, incy);
    if (alpha == 0.0) {
        if (beta == 1.0) return 0;
        int blocks = (leny + 255) / 256;
        if (blocks > 4096) blocks = 4096;
        hipLaunchKernelGGL(scal_vec_kernel, dim3(blocks), dim3(256), 0, s, leny, beta, d_y,
                           (long)incy, y0);
        return (int)hipGetLastError();
    }
    if (nt) {
        // ~2 workgroups per CU, each panel at least 16 columns wide
        const int nrb = ceil_div(m, GN_ROWS);
        int ncb = ceil_div(512, nrb);
        const int maxcb = ceil_div(n, 16);
        if (ncb > maxcb) ncb = maxcb;
        if (ncb < 1) ncb = 1;
        const int cbw = ceil_div(n, ncb);
        ncb = ceil_div(n, cbw);
        double *partials = nullptr;
        if (ncb > 1) {
            partials = (double *)dev_workspace((size_t)ncb * m * sizeof(double));
            if (!partials) return (int)hipErrorOutOfMemory;
        }
        const int vec = (((uintptr_t)d_A & 15) == 0) && (lda % 2 == 0);
        PROF_BEGIN(s, "dgemv_n");
        hipLaunchKernelGGL(gemv_n_stream, dim3((unsigned)nrb * ncb), dim3(256), 0, s, m, n, alpha,
                           d_A, (long)lda, d_x, (long)incx, x0, beta, d_y, (long)incy, y0, nrb,
                           cbw, ncb == 1, vec, partials);
        PROF_END(s);
        if (ncb > 1) {
            PROF_BEGIN(s, "dgemv_n_fold");
            hipLaunchKernelGGL(gemv_n_fold, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, s, m,
                               ncb, partials, alpha, beta, d_y, (long)incy, y0);
            PROF_END(s);
        }
        return (int)hipGetLastError();
    }
    const int vec = (((uintptr_t)d_A & 15) == 0) && (lda % 2 == 0) && (((uintptr_t)d_x & 15) == 0);
    int blocks = ceil_div(n, 4);
    if (blocks > 2048) blocks = 2048;
    PROF_BEGIN(s, "dgemv_t");
    hipLaunchKernelGGL(gemv_t_kernel, dim3((unsigned)blocks), dim3(256), 0, s, m, n, alpha, d_A,
                       (long)lda, d_x, (long)incx, x0, beta, d_y, (long)incy, y0, vec);
    PROF_END(s);
    return (int)hipGetLastError();
}

int gpublas_dev_dger(void *stream, int m, int n, double alpha, const double *d_x, int incx,
                     const double *d_y, int incy, double *d_A, int lda)
{
    hipStream_t s = pick_stream(stream);
    if (m <= 0 || n <= 0 || alpha == 0.0) return 0;
    dim3 grid((unsigned)ceil_div(m, 256), (unsigned)ceil_div(n, 16));
    PROF_BEGIN(s, "dger");
    hipLaunchKernelGGL(ger_kernel, grid, dim3(256), 0, s, m, n, alpha, d_x, (long)incx,
                       start_of(m, incx), d_y, (long)incy, start_of(n, incy), d_A, (long)lda);
    PROF_END(s);
    return (int)hipGetLastError();
}

int gpublas_dev_dtrsv(void *stream, char uplo, char trans, char diag, int n, const double *d_A,
                      int lda, double *d_x, int incx)
{
    hipStream_t s = pick_stream(stream);
    if (n <= 0) return 0;
    if (incx == 1) return gpublas_dev_dtrsm(s, 'L', uplo, trans, diag, n, 1, 1.0, d_A, lda, d_x, n);
    double *t = (double *)gpublas_dev_alloc((size_t)n * sizeof(double));
    if (!t) return (int)hipErrorOutOfMemory;
    const long x0 = start_of(n, incx);
    int blocks = ceil_div(n, 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(gather_kernel, dim3(blocks), dim3(256), 0, s, n, d_x, (long)incx, x0, t);
    int err = gpublas_dev_dtrsm(s, 'L', uplo, trans, diag, n, 1, 1.0, d_A, lda, t, n);
    if (!err) {
        hipLaunchKernelGGL(scatter_kernel, dim3(blocks), dim3(256), 0, s, n, t, d_x, (long)incx, x0);
        err = (int)hipGetLastError();
    }
    (void)hipStreamSynchronize(s);
    gpublas_dev_release(t);
    return err;
}

}  // extern "C"
