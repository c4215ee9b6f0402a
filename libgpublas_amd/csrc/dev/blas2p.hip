// blas2p.hip -- DTPSV on gfx950 (replaces cublasDtpsv behind the reference's
// unbuilt CBLAS wrapper, blas_level2/tpsv.cc).
//
// Packed storage (netlib column packing: column j of an upper triangle holds
// rows 0..j at AP[j(j+1)/2 ..], of a lower one rows j..n-1 at
// AP[j(2n-j+1)/2 ..]).  DTPSV unpacks it into a full-storage block of a pool
// buffer and runs the DTRSV solve.  The packed products DSPMV and DTPMV, and
// DTRMV, run the DSYMV tile kernel (blas2.hip), which reads the triangle once
// in place.
#include "devcommon.hpp"
#include "gpublas_dev.h"

namespace {

inline bool up(char c) { return c == 'U' || c == 'u'; }

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

}  // extern "C"
