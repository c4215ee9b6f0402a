// blas2p.hip -- triangular and packed Level-2 routines on gfx950: DTRMV,
// DTPMV, DTPSV, DSPMV.
//
// Replace the vendor calls of the reference's unbuilt CBLAS wrappers
// (cublasDtrmv, blas_level2/trmv.cc; cublasDtpmv / cublasDtpsv, tpmv.cc,
// tpsv.cc; cublasDspmv, pmv.cc).
//
// DTRMV x := op(T) x keeps a copy t of x (gathered through incx) and writes
// x from it:
//  * op = N: a 256-row x 256-column grid over the stored triangle; thread i of
//    a workgroup sums T(i, j) t(j) over the chunk's columns (one coalesced 2 KB
//    column read per step across the workgroup, t in LDS) into a partial per
//    column chunk; trmv_fold adds the row's chunks in order (deterministic).
//    (64-column chunks, 4x the workgroups, measured 20 % slower at n = 8192.)
//  * op = T: one wave per column j, a dot product over the column's stored
//    segment (contiguous) with a fixed shuffle tree.
// The packed routines unpack the triangle into a full-storage block of a
// pool buffer (netlib column packing: column j of an upper triangle holds
// rows 0..j at AP[j(j+1)/2 ..], of a lower one rows j..n-1 at
// AP[j(2n-j+1)/2 ..]) and run DTRMV, DTRSV or DSYMV on it: two more passes
// over the triangle than a packed kernel, in exchange for the tuned
// full-storage kernels.
#include "devcommon.hpp"
#include "gpublas_dev.h"

namespace {

constexpr int TR_ROWS = 256, TR_COLS = 256;

inline long start_of(int n, int inc) { return inc < 0 ? -(long)(n - 1) * inc : 0; }
inline bool up(char c) { return c == 'U' || c == 'u'; }
inline bool isN(char c) { return c == 'N' || c == 'n'; }

__global__ void gather_vec(int n, const double *x, long incx, long x0, double *t)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (long)gridDim.x * blockDim.x) t[i] = x[x0 + i * incx];
}

template <bool LOWER, bool UNIT>
__global__ __launch_bounds__(256) void trmv_n_part(int n, const double *A, long lda,
                                                   const double *t, double *P)
{
    const int rb = blockIdx.x, cb = blockIdx.y;
    // chunks wholly outside the triangle of this row block
    if (LOWER ? cb * TR_COLS > rb * TR_ROWS + TR_ROWS - 1 : cb * TR_COLS + TR_COLS - 1 < rb * TR_ROWS)
        return;
    __shared__ double ts[TR_COLS];
    const long j0 = (long)cb * TR_COLS;
    const int jn = (int)min((long)TR_COLS, n - j0);
    if (threadIdx.x < jn) ts[threadIdx.x] = t[j0 + threadIdx.x];
    __syncthreads();
    const long i = (long)rb * TR_ROWS + threadIdx.x;
    if (i >= n) return;
    // this row's columns in the chunk: lower j <= i, upper j >= i
    const int lo = LOWER ? 0 : (int)max(0L, i - j0), hi = LOWER ? (int)min((long)jn, i - j0 + 1) : jn;
    double s = 0;
    const double *a = A + i + j0 * lda;
    int j = lo;
    for (; j + 4 <= hi; j += 4) {
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = a[(long)(j + u) * lda];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            s = fma(UNIT && j0 + j + u == i ? 1.0 : v[u], ts[j + u], s);
    }
    for (; j < hi; ++j) s = fma(UNIT && j0 + j == i ? 1.0 : a[(long)j * lda], ts[j], s);
    P[(long)cb * n + i] = s;
}

template <bool LOWER>
__global__ __launch_bounds__(256) void trmv_fold(int n, int nc, const double *P, double *x,
                                                 long incx, long x0)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int b = (int)(i / TR_ROWS);
    // chunks feeding row i: lower 0 .. its diagonal chunk, upper from it on
    const int c_lo = LOWER ? 0 : (int)((long)b * TR_ROWS / TR_COLS);
    const int c_hi = LOWER ? (int)min((long)nc - 1, ((long)b * TR_ROWS + TR_ROWS - 1) / TR_COLS) : nc - 1;
    double s = 0;
    for (int c = c_lo; c <= c_hi; ++c) s += P[(long)c * n + i];
    x[x0 + i * incx] = s;
}

template <bool LOWER, bool UNIT>
__global__ __launch_bounds__(256) void trmv_t_kernel(int n, const double *A, long lda,
                                                     const double *t, double *x, long incx,
                                                     long x0)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (long j = (long)blockIdx.x * 4 + wave; j < n; j += (long)gridDim.x * 4) {
        const double *a = A + j * lda;
        const long lo = LOWER ? j : 0, hi = LOWER ? n : j + 1;
        // four loads of A and t in flight per lane; two accumulators
        double s = 0, s2 = 0;
        long i = lo + lane;
        for (; i + 192 < hi; i += 256) {
            double av[4], tv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                av[u] = a[i + 64 * u];
                tv[u] = t[i + 64 * u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const double v = UNIT && i + 64 * u == j ? 1.0 : av[u];
                if (u & 1) s2 = fma(v, tv[u], s2);
                else s = fma(v, tv[u], s);
            }
        }
        for (; i < hi; i += 64) s = fma(UNIT && i == j ? 1.0 : a[i], t[i], s);
        s = wave_sum(s + s2);
        if (lane == 0) x[x0 + j * incx] = s;
    }
}

// AP (packed) -> W (full storage, leading dimension n), stored triangle only.
template <bool LOWER>
__global__ __launch_bounds__(256) void unpack_kernel(int n, const double *AP, double *W)
{
    const long r0 = (long)blockIdx.x * 256, c0 = (long)blockIdx.y * 16;
    if (LOWER ? r0 + 255 < c0 : r0 > c0 + 15) return;
    const long r = r0 + threadIdx.x;
    if (r >= n) return;
    for (int u = 0; u < 16; ++u) {
        const long j = c0 + u;
        if (j >= n || (LOWER ? r < j : r > j)) continue;
        const long k = LOWER ? j * (2L * n - j + 1) / 2 + (r - j) : j * (j + 1) / 2 + r;
        W[r + j * (long)n] = AP[k];
    }
}

int trmv(hipStream_t s, char uplo, char trans, char diag, int n, const double *A, long lda,
         double *x, int incx)
{
    if (n <= 0) return 0;
    const bool lower = !up(uplo), unit = !isN(diag);
    double *t = (double *)gpublas_dev_alloc((size_t)n * sizeof(double));
    if (!t) return (int)hipErrorOutOfMemory;
    const long x0 = start_of(n, incx);
    hipLaunchKernelGGL(gather_vec, dim3((unsigned)std::min(ceil_div(n, 256), 4096)), dim3(256), 0,
                       s, n, x, (long)incx, x0, t);
    PROF_BEGIN(s, "dtrmv");
    if (isN(trans)) {
        const int nr = ceil_div(n, TR_ROWS), nc = ceil_div(n, TR_COLS);
        double *P = (double *)gpublas_dev_alloc((size_t)nc * n * sizeof(double));
        if (!P) {
            dev_release_after(t, s);
            return (int)hipErrorOutOfMemory;
        }
        dim3 grid((unsigned)nr, (unsigned)nc);
#define GB_TRN(L, U) hipLaunchKernelGGL((trmv_n_part<L, U>), grid, dim3(256), 0, s, n, A, lda, t, P)
        if (lower) { if (unit) GB_TRN(true, true); else GB_TRN(true, false); }
        else { if (unit) GB_TRN(false, true); else GB_TRN(false, false); }
#undef GB_TRN
        if (lower)
            hipLaunchKernelGGL(trmv_fold<true>, dim3((unsigned)nr), dim3(256), 0, s, n, nc, P, x,
                               (long)incx, x0);
        else
            hipLaunchKernelGGL(trmv_fold<false>, dim3((unsigned)nr), dim3(256), 0, s, n, nc, P, x,
                               (long)incx, x0);
        dev_release_after(P, s);
    } else {
        const unsigned blocks = (unsigned)std::min(ceil_div(n, 4), 8192);
#define GB_TRT(L, U)                                                                              \
    hipLaunchKernelGGL((trmv_t_kernel<L, U>), dim3(blocks), dim3(256), 0, s, n, A, lda, t, x,      \
                       (long)incx, x0)
        if (lower) { if (unit) GB_TRT(true, true); else GB_TRT(true, false); }
        else { if (unit) GB_TRT(false, true); else GB_TRT(false, false); }
#undef GB_TRT
    }
    PROF_END(s);
    dev_release_after(t, s);
    return (int)hipGetLastError();
}

// unpack AP into a pool block; the caller releases it after its last use
double *unpack(hipStream_t s, char uplo, int n, const double *AP)
{
    double *W = (double *)gpublas_dev_alloc((size_t)n * n * sizeof(double));
    if (!W) return nullptr;
    dim3 grid((unsigned)ceil_div(n, 256), (unsigned)ceil_div(n, 16));
    if (up(uplo)) hipLaunchKernelGGL(unpack_kernel<false>, grid, dim3(256), 0, s, n, AP, W);
    else hipLaunchKernelGGL(unpack_kernel<true>, grid, dim3(256), 0, s, n, AP, W);
    return W;
}

}  // namespace

extern "C" {

int gpublas_dev_dtrmv(void *stream, char uplo, char trans, char diag, int n, const double *d_A,
                      int lda, double *d_x, int incx)
{
    return trmv(pick_stream(stream), uplo, trans, diag, n, d_A, lda, d_x, incx);
}

int gpublas_dev_dtpmv(void *stream, char uplo, char trans, char diag, int n, const double *d_AP,
                      double *d_x, int incx)
{
    hipStream_t s = pick_stream(stream);
    if (n <= 0) return 0;
    double *W = unpack(s, uplo, n, d_AP);
    if (!W) return (int)hipErrorOutOfMemory;
    const int e = trmv(s, uplo, trans, diag, n, W, n, d_x, incx);
    dev_release_after(W, s);
    return e;
}

int gpublas_dev_dtpsv(void *stream, char uplo, char trans, char diag, int n, const double *d_AP,
                      double *d_x, int incx)
{
    hipStream_t s = pick_stream(stream);
    if (n <= 0) return 0;
    double *W = unpack(s, uplo, n, d_AP);
    if (!W) return (int)hipErrorOutOfMemory;
    const int e = gpublas_dev_dtrsv(s, uplo, trans, diag, n, W, n, d_x, incx);
    dev_release_after(W, s);
    return e;
}

int gpublas_dev_dspmv(void *stream, char uplo, int n, double alpha, const double *d_AP,
                      const double *d_x, int incx, double beta, double *d_y, int incy)
{
    hipStream_t s = pick_stream(stream);
    if (n <= 0) return 0;
    if (alpha == 0.0) return gpublas_dev_dsymv(s, uplo, n, alpha, nullptr, n, d_x, incx, beta, d_y, incy);
    double *W = unpack(s, uplo, n, d_AP);
    if (!W) return (int)hipErrorOutOfMemory;
    const int e = gpublas_dev_dsymv(s, uplo, n, alpha, W, n, d_x, incx, beta, d_y, incy);
    dev_release_after(W, s);
    return e;
}

}  // extern "C"
