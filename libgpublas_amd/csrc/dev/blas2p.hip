// blas2p.hip -- triangular and packed Level-2 routines on gfx950: DTRMV,
// DTPMV, DTPSV (DSPMV runs the DSYMV tile kernel on packed storage, blas2.hip).
//
// Replace the vendor calls of the reference's unbuilt CBLAS wrappers
// (cublasDtrmv, blas_level2/trmv.cc; cublasDtpmv / cublasDtpsv, tpmv.cc,
// tpsv.cc; cublasDspmv, pmv.cc).
//
// DTRMV x := op(T) x:
//  * op = N: the streaming DGEMV layout (512-row x cbw-column panels, 16-B
//    non-temporal loads, eight columns in flight per wave) with each panel
//    clipped to the stored triangle, reading x in place; partial sums per
//    panel, folded in panel order by trmv_fold (deterministic), which then
//    writes x.  (A first version with 256 x 256
//    chunks and 8-B loads ran at 3.05 TB/s at n = 8192.)
//  * op = T: one wave per column j, a dot product over the column's stored
//    segment (contiguous) with a fixed shuffle tree, from a copy t of x
//    (gathered through incx) since x(j) is written while others read.
// Packed storage (netlib column packing: column j of an upper triangle holds
// rows 0..j at AP[j(j+1)/2 ..], of a lower one rows j..n-1 at
// AP[j(2n-j+1)/2 ..]): DTPMV reads it in place (pk_n_stream, pk_t_kernel);
// DTPSV unpacks it into a full-storage block of a pool buffer and runs the
// DTRSV solve.
#include "devcommon.hpp"
#include "gpublas_dev.h"

namespace {


inline long start_of(int n, int inc) { return inc < 0 ? -(long)(n - 1) * inc : 0; }
inline bool up(char c) { return c == 'U' || c == 'u'; }
inline bool isN(char c) { return c == 'N' || c == 'n'; }

__global__ void gather_vec(int n, const double *x, long incx, long x0, double *t)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (long)gridDim.x * blockDim.x) t[i] = x[x0 + i * incx];
}

typedef double d2v __attribute__((ext_vector_type(2)));
constexpr int TS_ROWS = 512, TS_U = 8;

// op = N: the streaming DGEMV layout on the stored triangle.  A workgroup owns
// a 512-row x cbw-column panel clipped to the triangle; wave w reads rows
// [128w, 128w + 128) of each column, 16 B per lane, eight columns in flight.
// Columns wholly inside the triangle for the panel's rows take the unmasked
// loop; in the 512 columns of the diagonal block the elements outside the
// triangle are selected away (a unit diagonal is taken as one).
template <bool LOWER, bool UNIT>
__global__ __launch_bounds__(256) void trmv_n_stream(int n, const double *A, long lda,
                                                     const double *x, long incx, long x0, int nrb,
                                                     int cbw, int vec, double *P)
{
    // x is only read here; trmv_fold, the next launch on the stream, writes it
    auto t = [&](long c) { return x[x0 + c * incx]; };
    const int rb = blockIdx.x % nrb, cb = blockIdx.x / nrb;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long r0 = (long)rb * TS_ROWS, row = r0 + wave * 128 + 2 * lane;
    long c0 = (long)cb * cbw, c1 = min((long)n, c0 + cbw);
    // stored columns of this row block: lower [0, r0 + 512), upper [r0, n)
    if (LOWER) c1 = min(c1, r0 + TS_ROWS);
    else c0 = max(c0, r0);
    if (c0 >= c1) return;  // the fold reads no partials of this panel
    // unmasked columns: lower j < r0, upper j >= r0 + 512
    const long f0 = LOWER ? c0 : max(c0, r0 + TS_ROWS), f1 = LOWER ? min(c1, r0) : c1;
    const long d0 = max(c0, r0), d1 = min(c1, r0 + TS_ROWS);
    double s0 = 0, s1 = 0;
    if (f0 < f1) {
        if (vec && row + 1 < n) {
            const double *a = A + row;
            long c = f0;
            for (; c + TS_U <= f1; c += TS_U) {
                d2v v[TS_U];
#pragma unroll
                for (int u = 0; u < TS_U; ++u)
                    v[u] = __builtin_nontemporal_load((const d2v *)(a + (c + u) * lda));
#pragma unroll
                for (int u = 0; u < TS_U; ++u) {
                    const double xv = t(c + u);
                    s0 = fma(v[u].x, xv, s0);
                    s1 = fma(v[u].y, xv, s1);
                }
            }
            for (; c < f1; ++c) {
                const d2v v = *(const d2v *)(a + c * lda);
                s0 = fma(v.x, t(c), s0);
                s1 = fma(v.y, t(c), s1);
            }
        } else {
            for (long c = f0; c < f1; ++c) {
                if (row < n) s0 = fma(A[row + c * lda], t(c), s0);
                if (row + 1 < n) s1 = fma(A[row + 1 + c * lda], t(c), s1);
            }
        }
    }
    // diagonal block: the same loads (the device operand holds the full
    // lda x n array, so the other triangle's memory is in bounds), with the
    // elements outside the triangle replaced by zero -- selected, never
    // multiplied, so whatever they hold cannot reach the sum
    auto keep = [&](long i, long c, double v) {
        return (LOWER ? i >= c : i <= c) ? (UNIT && i == c ? 1.0 : v) : 0.0;
    };
    if (d0 < d1) {
        if (vec && row + 1 < n) {
            const double *a = A + row;
            long c = d0;
            for (; c + TS_U <= d1; c += TS_U) {
                d2v v[TS_U];
#pragma unroll
                for (int u = 0; u < TS_U; ++u) v[u] = *(const d2v *)(a + (c + u) * lda);
#pragma unroll
                for (int u = 0; u < TS_U; ++u) {
                    const double xv = t(c + u);
                    s0 = fma(keep(row, c + u, v[u].x), xv, s0);
                    s1 = fma(keep(row + 1, c + u, v[u].y), xv, s1);
                }
            }
            for (; c < d1; ++c) {
                const d2v v = *(const d2v *)(a + c * lda);
                s0 = fma(keep(row, c, v.x), t(c), s0);
                s1 = fma(keep(row + 1, c, v.y), t(c), s1);
            }
        } else {
            for (long c = d0; c < d1; ++c) {
                if (row < n) s0 = fma(keep(row, c, A[row + c * lda]), t(c), s0);
                if (row + 1 < n) s1 = fma(keep(row + 1, c, A[row + 1 + c * lda]), t(c), s1);
            }
        }
    }
    double *pp = P + (long)cb * n;
    if (row < n) pp[row] = s0;
    if (row + 1 < n) pp[row + 1] = s1;
}

// x(i) := sum of the row's panel partials, in panel order (deterministic)
template <bool LOWER>
__global__ __launch_bounds__(64) void trmv_fold(int n, int ncb, int cbw, const double *P,
                                                double *x, long incx, long x0)
{
    const long i = (long)blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    const long r0 = i / TS_ROWS * TS_ROWS;
    // the panels with stored columns for this row block (each wrote its rows)
    const int c_lo = LOWER ? 0 : (int)(r0 / cbw);
    const int c_hi = LOWER ? (int)min((long)ncb - 1, (r0 + TS_ROWS - 1) / cbw) : ncb - 1;
    // eight loads in flight, added in panel order
    double s = 0;
    int c = c_lo;
    for (; c + 8 <= c_hi + 1; c += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = P[(long)(c + u) * n + i];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; c <= c_hi; ++c) s += P[(long)c * n + i];
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

// ---- packed storage read in place (DTPMV) -----------------------------------
// Column j of the packed triangle is contiguous: lower rows j .. n-1 from
// AP[j (2n - j + 1) / 2], upper rows 0 .. j from AP[j (j + 1) / 2].  pcol(j)
// is the address of element (0, j) of that column, so pcol(j)[i] is (i, j)
// for every stored i.  Inside the diagonal block the 'N' kernel also reads
// the few entries of the neighbouring column next to the stored run (rows on
// the other side of the diagonal); those addresses stay inside AP and the
// values are selected away.
template <bool LOWER>
__device__ __forceinline__ const double *pcol(const double *AP, long n, long j)
{
    return LOWER ? AP + j * (2 * n - j + 1) / 2 - j : AP + j * (j + 1) / 2;
}

// op = N on the packed triangle: the panel layout of trmv_n_stream with 8-B
// loads (a packed column's 16-B alignment alternates), two rows per lane,
// eight columns in flight.
template <bool LOWER, bool UNIT>
__global__ __launch_bounds__(256) void pk_n_stream(int n, const double *AP, const double *x,
                                                   long incx, long x0, int nrb, int cbw, double *P)
{
    const int rb = blockIdx.x % nrb, cb = blockIdx.x / nrb;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long r0 = (long)rb * TS_ROWS, row = r0 + wave * 128 + 2 * lane;
    long c0 = (long)cb * cbw, c1 = min((long)n, c0 + cbw);
    if (LOWER) c1 = min(c1, r0 + TS_ROWS);
    else c0 = max(c0, r0);
    if (c0 >= c1) return;
    auto keep = [&](long i, long c, double v) {
        return (LOWER ? i >= c : i <= c) ? (UNIT && i == c ? 1.0 : v) : 0.0;
    };
    // loads are unconditional: for 0 <= c, row < n, pcol(c) + row is
    // AP + c (2n - c - 1) / 2 + row (lower) or AP + c (c + 1) / 2 + row
    // (upper), both within AP[0 .. n (n + 1) / 2 - 1]; the select in keep()
    // drops what lies outside the triangle.  row1 repeats row on the
    // matrix's last row (its value is not stored then).
    const bool r1ok = row + 1 < n;
    const long row1 = r1ok ? row + 1 : row;
    double s0 = 0, s1 = 0;
    if (row < n) {
        long c = c0;
        for (; c + TS_U <= c1; c += TS_U) {
            double v0[TS_U], v1[TS_U], xv[TS_U];
#pragma unroll
            for (int u = 0; u < TS_U; ++u) {
                const double *col = pcol<LOWER>(AP, n, c + u);
                v0[u] = col[row];
                v1[u] = col[row1];
                xv[u] = x[x0 + (c + u) * incx];
            }
#pragma unroll
            for (int u = 0; u < TS_U; ++u) {
                s0 = fma(keep(row, c + u, v0[u]), xv[u], s0);
                s1 = fma(keep(row1, c + u, v1[u]), xv[u], s1);
            }
        }
        for (; c < c1; ++c) {
            const double *col = pcol<LOWER>(AP, n, c);
            const double xv = x[x0 + c * incx];
            s0 = fma(keep(row, c, col[row]), xv, s0);
            s1 = fma(keep(row1, c, col[row1]), xv, s1);
        }
    }
    double *pp = P + (long)cb * n;
    if (row < n) pp[row] = s0;
    if (r1ok) pp[row + 1] = s1;
}

// op = T on the packed triangle: one wave per column, a dot product over the
// column's contiguous stored run (the stored diagonal when DIAG == 1, a unit
// one when DIAG == 2), four loads in flight per lane.
template <bool LOWER, int DIAG>
__global__ __launch_bounds__(256) void pk_t_kernel(int n, const double *AP, const double *t,
                                                   long inct, long t0, double *out, long inco,
                                                   long o0)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (long j = (long)blockIdx.x * 4 + wave; j < n; j += (long)gridDim.x * 4) {
        const double *a = pcol<LOWER>(AP, n, j);
        const long lo = LOWER ? j + (DIAG != 1) : 0, hi = LOWER ? n : j + (DIAG == 1);
        double s = 0, s2 = 0;
        long i = lo + lane;
        for (; i + 192 < hi; i += 256) {
            double av[4], tv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                av[u] = a[i + 64 * u];
                tv[u] = t[t0 + (i + 64 * u) * inct];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (u & 1) s2 = fma(av[u], tv[u], s2);
                else s = fma(av[u], tv[u], s);
            }
        }
        for (; i < hi; i += 64) s = fma(a[i], t[t0 + i * inct], s);
        s = wave_sum(s + s2);
        if (lane == 0) out[o0 + j * inco] = DIAG == 2 ? s + t[t0 + j * inct] : s;
    }
}

// out(i) := alpha * (panel partials of row i in panel order) + beta * out(i)
template <bool LOWER>
__global__ __launch_bounds__(64) void pk_fold(int n, int ncb, int cbw, const double *P,
                                              double alpha, double beta, double *y, long incy,
                                              long y0)
{
    const long i = (long)blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    const long r0 = i / TS_ROWS * TS_ROWS;
    const int c_lo = LOWER ? 0 : (int)(r0 / cbw);
    const int c_hi = LOWER ? (int)min((long)ncb - 1, (r0 + TS_ROWS - 1) / cbw) : ncb - 1;
    double s = 0;
    int c = c_lo;
    for (; c + 8 <= c_hi + 1; c += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = P[(long)(c + u) * n + i];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; c <= c_hi; ++c) s += P[(long)c * n + i];
    double *yo = y + y0 + i * incy;
    *yo = beta == 0.0 ? alpha * s : fma(alpha, s, beta * *yo);
}

// op(T) x on packed storage (DTPMV).
int packed_mv(hipStream_t s, bool lower, bool trans, bool unit, int n, const double *AP,
              const double *x, int incx, double alpha, double beta, double *y, int incy)
{
    const long x0 = start_of(n, incx), y0 = start_of(n, incy);
    const unsigned tblocks = (unsigned)std::min(ceil_div(n, 4), 8192);
    if (trans) {
        // x(j) is written while other waves read x(i): work from a copy
        double *t = (double *)gpublas_dev_alloc((size_t)n * sizeof(double));
        if (!t) return (int)hipErrorOutOfMemory;
        hipLaunchKernelGGL(gather_vec, dim3((unsigned)std::min(ceil_div(n, 256), 4096)), dim3(256),
                           0, s, n, x, (long)incx, x0, t);
        PROF_BEGIN(s, "dtpmv");
#define GB_PT(L, D) \
    hipLaunchKernelGGL((pk_t_kernel<L, D>), dim3(tblocks), dim3(256), 0, s, n, AP, t, 1L, 0L, y, \
                       (long)incy, y0)
        if (lower) { if (unit) GB_PT(true, 2); else GB_PT(true, 1); }
        else { if (unit) GB_PT(false, 2); else GB_PT(false, 1); }
#undef GB_PT
        PROF_END(s);
        dev_release_after(t, s);
        return (int)hipGetLastError();
    }
    const int nrb = ceil_div(n, TS_ROWS);
    int ncb = ceil_div(1024, nrb);
    const int maxcb = ceil_div(n, 16);
    if (ncb > maxcb) ncb = maxcb;
    const int cbw = ceil_div(n, ncb);
    ncb = ceil_div(n, cbw);
    double *P = (double *)gpublas_dev_alloc((size_t)ncb * n * sizeof(double));
    if (!P) return (int)hipErrorOutOfMemory;
    PROF_BEGIN(s, "dtpmv");
    const dim3 grid((unsigned)nrb * ncb);
#define GB_PN(L, U) \
    hipLaunchKernelGGL((pk_n_stream<L, U>), grid, dim3(256), 0, s, n, AP, x, (long)incx, x0, nrb, cbw, P)
    if (lower) { if (unit) GB_PN(true, true); else GB_PN(true, false); }
    else { if (unit) GB_PN(false, true); else GB_PN(false, false); }
#undef GB_PN
    const dim3 fg((unsigned)ceil_div(n, 64));
#define GB_PF(L) \
    hipLaunchKernelGGL((pk_fold<L>), fg, dim3(64), 0, s, n, ncb, cbw, P, alpha, beta, y, \
                       (long)incy, y0)
    if (lower) GB_PF(true);
    else GB_PF(false);
#undef GB_PF
    PROF_END(s);
    dev_release_after(P, s);
    return (int)hipGetLastError();
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
    const long x0 = start_of(n, incx);
    if (isN(trans)) {
        PROF_BEGIN(s, "dtrmv");
        // ~1024 panels of at least 16 columns, about half in the triangle
        // (2048 measured 3 % slower at n = 8192 and 4 % faster at 30000)
        const int nrb = ceil_div(n, TS_ROWS);
        int ncb = ceil_div(1024, nrb);
        const int maxcb = ceil_div(n, 16);
        if (ncb > maxcb) ncb = maxcb;
        const int cbw = ceil_div(n, ncb);
        ncb = ceil_div(n, cbw);
        double *P = (double *)gpublas_dev_alloc((size_t)ncb * n * sizeof(double));
        if (!P) {
            PROF_END(s);
            return (int)hipErrorOutOfMemory;
        }
        const int vec = (((uintptr_t)A & 15) == 0) && (lda % 2 == 0);
        const dim3 grid((unsigned)nrb * ncb);
#define GB_TRN(L, U)                                                                              \
    hipLaunchKernelGGL((trmv_n_stream<L, U>), grid, dim3(256), 0, s, n, A, lda, x, (long)incx, x0, \
                       nrb, cbw, vec, P)
        if (lower) { if (unit) GB_TRN(true, true); else GB_TRN(true, false); }
        else { if (unit) GB_TRN(false, true); else GB_TRN(false, false); }
#undef GB_TRN
        if (lower)
            hipLaunchKernelGGL(trmv_fold<true>, dim3((unsigned)ceil_div(n, 64)), dim3(64), 0, s, n,
                               ncb, cbw, P, x, (long)incx, x0);
        else
            hipLaunchKernelGGL(trmv_fold<false>, dim3((unsigned)ceil_div(n, 64)), dim3(64), 0, s, n,
                               ncb, cbw, P, x, (long)incx, x0);
        dev_release_after(P, s);
        PROF_END(s);
    } else {
        // op = T writes x(j) while other waves still read x(i): work from a copy
        double *t = (double *)gpublas_dev_alloc((size_t)n * sizeof(double));
        if (!t) return (int)hipErrorOutOfMemory;
        hipLaunchKernelGGL(gather_vec, dim3((unsigned)std::min(ceil_div(n, 256), 4096)), dim3(256),
                           0, s, n, x, (long)incx, x0, t);
        PROF_BEGIN(s, "dtrmv");
        const unsigned blocks = (unsigned)std::min(ceil_div(n, 4), 8192);
#define GB_TRT(L, U)                                                                              \
    hipLaunchKernelGGL((trmv_t_kernel<L, U>), dim3(blocks), dim3(256), 0, s, n, A, lda, t, x,      \
                       (long)incx, x0)
        if (lower) { if (unit) GB_TRT(true, true); else GB_TRT(true, false); }
        else { if (unit) GB_TRT(false, true); else GB_TRT(false, false); }
#undef GB_TRT
        PROF_END(s);
        dev_release_after(t, s);
    }
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
    return packed_mv(s, !up(uplo), !isN(trans), !isN(diag), n, d_AP, d_x, incx, 1.0, 0.0,
                     d_x, incx);
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


}  // extern "C"
