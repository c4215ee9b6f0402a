// blas2.hip -- Level-2 kernels on gfx950: DGEMV (N and T), DGER, DTRSV.
//
// Replaces the vendor calls of the reference's unbuilt CBLAS wrappers:
// cublasDgemv (blas_level2/gemv.cc:97), cublasDger (ger.cc:92),
// cublasDtrsv (trsv.cc:86).
//
// DGEMV 'N' is the HBM stream of the Level-1/2 chain: y = alpha*A*x + beta*y
// with A column-major.  A workgroup owns 128 rows x CB columns; each of its
// 4 waves reads whole 1 KB column segments (64 lanes x 16 B) of every 4th
// column, so each load instruction is one contiguous 1 KB burst.  The waves'
// partial sums are folded through LDS in wave order; when the columns are
// split over several workgroups, each writes its 128 partial sums and the
// workgroup drawing the last ticket of that row block folds the column
// blocks in index order and applies alpha / beta.  The reduction order is a
// function of (m, n) only, so results are reproducible.
// DGEMV 'T' runs one wave per output element (a column of A), 16-B loads and
// a wave64 shuffle tree.
#include "devcommon.hpp"
#include "gpublas_dev.h"

int gpublas_dev_dtrsm(void *stream, char side, char uplo, char transa, char diag, int m, int n,
                      double alpha, const double *d_A, int lda, double *d_B, int ldb);

namespace {

constexpr int GEMV_ROWS = 128;

__global__ __launch_bounds__(256) void gemv_n_kernel(int m, int n, double alpha, const double *A,
                                                     long lda, const double *x, long incx, long x0,
                                                     double beta, double *y, long incy, long y0,
                                                     int nrb, int ncb, int cbw, int vec,
                                                     double *partials, unsigned *tickets)
{
    __shared__ double red[4][GEMV_ROWS];
    __shared__ bool is_last;
    const int rb = blockIdx.x % nrb, cb = blockIdx.x / nrb;
    const int r0 = rb * GEMV_ROWS;
    const int c0 = cb * cbw;
    const int c1 = min(n, c0 + cbw);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int row = r0 + 2 * lane;
    double s0 = 0, s1 = 0;
    if (vec && row + 1 < m) {
        const double *a = A + row;
        int c = c0 + wave;
        for (; c + 12 < c1; c += 16) {
            const double2 v0 = *(const double2 *)(a + (long)c * lda);
            const double2 v1 = *(const double2 *)(a + (long)(c + 4) * lda);
            const double2 v2 = *(const double2 *)(a + (long)(c + 8) * lda);
            const double2 v3 = *(const double2 *)(a + (long)(c + 12) * lda);
            const double x0v = x[x0 + (long)c * incx], x1v = x[x0 + (long)(c + 4) * incx];
            const double x2v = x[x0 + (long)(c + 8) * incx], x3v = x[x0 + (long)(c + 12) * incx];
            s0 = fma(v0.x, x0v, s0); s1 = fma(v0.y, x0v, s1);
            s0 = fma(v1.x, x1v, s0); s1 = fma(v1.y, x1v, s1);
            s0 = fma(v2.x, x2v, s0); s1 = fma(v2.y, x2v, s1);
            s0 = fma(v3.x, x3v, s0); s1 = fma(v3.y, x3v, s1);
        }
        for (; c < c1; c += 4) {
            const double2 v = *(const double2 *)(a + (long)c * lda);
            const double xv = x[x0 + (long)c * incx];
            s0 = fma(v.x, xv, s0);
            s1 = fma(v.y, xv, s1);
        }
    } else {
        for (int c = c0 + wave; c < c1; c += 4) {
            const double xv = x[x0 + (long)c * incx];
            if (row < m) s0 = fma(A[row + (long)c * lda], xv, s0);
            if (row + 1 < m) s1 = fma(A[row + 1 + (long)c * lda], xv, s1);
        }
    }
    red[wave][2 * lane] = s0;
    red[wave][2 * lane + 1] = s1;
    __syncthreads();
    const int t = threadIdx.x;
    const int i = r0 + t;
    double v = 0;
    if (t < GEMV_ROWS) v = red[0][t] + red[1][t] + red[2][t] + red[3][t];
    if (ncb == 1) {
        if (t < GEMV_ROWS && i < m) {
            double *yp = y + y0 + (long)i * incy;
            *yp = beta == 0.0 ? alpha * v : fma(beta, *yp, alpha * v);
        }
        return;
    }
    if (t < GEMV_ROWS && i < m) partials[(long)cb * m + i] = v;
    __syncthreads();
    if (t == 0) {
        __threadfence();
        unsigned k = __hip_atomic_fetch_add(&tickets[rb], 1u, __ATOMIC_ACQ_REL,
                                            __HIP_MEMORY_SCOPE_AGENT);
        is_last = (k == (unsigned)ncb - 1);
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    if (t < GEMV_ROWS && i < m) {
        double acc = 0;
        for (int b = 0; b < ncb; ++b) acc += partials[(long)b * m + i];
        double *yp = y + y0 + (long)i * incy;
        *yp = beta == 0.0 ? alpha * acc : fma(beta, *yp, alpha * acc);
    }
    if (t == 0) __hip_atomic_store(&tickets[rb], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(256) void gemv_t_kernel(int m, int n, double alpha, const double *A,
                                                     long lda, const double *x, long incx, long x0,
                                                     double beta, double *y, long incy, long y0,
                                                     int vec)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long j = (long)blockIdx.x * 4 + wave;
    if (j >= n) return;
    const double *a = A + j * lda;
    double s = 0;
    if (vec && incx == 1 && x0 == 0) {
        int i = 2 * lane;
        for (; i + 1 < m; i += 128) {
            const double2 av = *(const double2 *)(a + i);
            const double2 xv = *(const double2 *)(x + i);
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
        const int nrb = ceil_div(m, GEMV_ROWS);
        // Aim for ~4 workgroups per CU; each workgroup at least 64 columns.
        int ncb = ceil_div(1024, nrb);
        const int maxcb = ceil_div(n, 64);
        if (ncb > maxcb) ncb = maxcb;
        if (ncb < 1) ncb = 1;
        if (nrb > kTicketSlots) ncb = 1;
        const int cbw = ceil_div(n, ncb);
        ncb = ceil_div(n, cbw);
        double *partials = nullptr;
        if (ncb > 1) {
            partials = (double *)dev_workspace((size_t)ncb * m * sizeof(double));
            if (!partials) return (int)hipErrorOutOfMemory;
        }
        const int vec = (((uintptr_t)d_A & 15) == 0) && (lda % 2 == 0);
        PROF_BEGIN(s, "dgemv_n");
        hipLaunchKernelGGL(gemv_n_kernel, dim3((unsigned)(nrb * ncb)), dim3(256), 0, s, m, n,
                           alpha, d_A, (long)lda, d_x, (long)incx, x0, beta, d_y, (long)incy, y0,
                           nrb, ncb, cbw, vec, partials, dev_ctx()->tickets);
        PROF_END(s);
        return (int)hipGetLastError();
    }
    const int vec = (((uintptr_t)d_A & 15) == 0) && (lda % 2 == 0) && (((uintptr_t)d_x & 15) == 0);
    PROF_BEGIN(s, "dgemv_t");
    hipLaunchKernelGGL(gemv_t_kernel, dim3((unsigned)ceil_div(n, 4)), dim3(256), 0, s, m, n, alpha,
                       d_A, (long)lda, d_x, (long)incx, x0, beta, d_y, (long)incy, y0, vec);
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
