// blas2.hip -- Level-2 kernels on gfx950: DGEMV (N and T), DGER, DSYMV,
// DSPMV, DTRMV, DTPMV (one tile kernel over a stored triangle, full or
// packed), DSYR, DSYR2, and the packed rank updates DSPR, DSPR2 (DTRSV:
// trsv.hip).
//
// Replaces the vendor calls of the reference's unbuilt CBLAS wrappers:
// cublasDgemv (blas_level2/gemv.cc:97), cublasDger (ger.cc:92).
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
    // all of a group's loads are issued before its sums (one memory round
    // trip per 16 panels); the sum runs in panel order
    constexpr int G = 16;
    double s = 0;
    int b = 0;
    for (; b + G <= ncb; b += G) {
        double p[G];
#pragma unroll
        for (int u = 0; u < G; ++u) p[u] = partials[(long)(b + u) * m + i];
#pragma unroll
        for (int u = 0; u < G; ++u) s += p[u];
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

// DGER: A := alpha x y^T + A, read and written once (16 mn bytes).  The grid
// is the DGEMV 'N' panel plan at one workgroup per CU (a read-modify-write
// stream measured 6.2 TB/s with one workgroup per CU, 5.7 with two; a sweep
// of 4 KB column runs by all workgroups together, as a plain copy does it,
// measured slower here): a workgroup owns GN_ROWS rows (two per lane,
// one 16-B load and store per column) of a panel of cbw columns and keeps
// GR_U columns of loads in flight per lane before their stores.  As netlib,
// a column with y(j) = 0 is left untouched and the update is
// A(i,j) + x(i) * (alpha y(j)) with one rounding (FMA).
constexpr int GR_U = 8;
__global__ __launch_bounds__(256) void ger_stream(int m, int n, double alpha, const double *x,
                                                  long incx, long x0, const double *y, long incy,
                                                  long y0, double *A, long lda, int nrb, int cbw,
                                                  int vec)
{
    const int rb = blockIdx.x % nrb, cb = blockIdx.x / nrb;
    const long row = (long)rb * GN_ROWS + 2 * threadIdx.x;
    const int c0 = cb * cbw, c1 = min(n, c0 + cbw);
    if (row >= m) return;
    const double xa = x[x0 + row * incx];
    const double xb = row + 1 < m ? x[x0 + (row + 1) * incx] : 0.0;
    if (vec && row + 1 < m) {
        double *a = A + row;
        int c = c0;
        for (; c + GR_U <= c1; c += GR_U) {
            d2 v[GR_U];
            double t[GR_U];
#pragma unroll
            for (int u = 0; u < GR_U; ++u) {
                v[u] = __builtin_nontemporal_load((const d2 *)(a + (long)(c + u) * lda));
                t[u] = y[y0 + (long)(c + u) * incy];
            }
#pragma unroll
            for (int u = 0; u < GR_U; ++u) {
                if (t[u] == 0.0) continue;
                const double ay = alpha * t[u];
                d2 w;
                w.x = fma(xa, ay, v[u].x);
                w.y = fma(xb, ay, v[u].y);
                __builtin_nontemporal_store(w, (d2 *)(a + (long)(c + u) * lda));
            }
        }
        for (; c < c1; ++c) {
            const double t = y[y0 + (long)c * incy];
            if (t == 0.0) continue;
            const double ay = alpha * t;
            d2 *p = (d2 *)(a + (long)c * lda);
            d2 w = *p;
            w.x = fma(xa, ay, w.x);
            w.y = fma(xb, ay, w.y);
            *p = w;
        }
        return;
    }
    for (int c = c0; c < c1; ++c) {
        const double t = y[y0 + (long)c * incy];
        if (t == 0.0) continue;
        const double ay = alpha * t;
        double *p = A + row + (long)c * lda;
        p[0] = fma(xa, ay, p[0]);
        if (row + 1 < m) p[1] = fma(xb, ay, p[1]);
    }
}

// DSYMV: y = alpha*A*x + beta*y with only the uplo triangle of A stored.
// Every stored element A(i,j), i != j, feeds two outputs: A(i,j)*x(j) into
// y(i) and A(i,j)*x(i) into y(j), so the kernel reads the triangle once
// (n^2/2 elements, half of a DGEMV).  The grid runs over the stored
// SY_T x SY_T tiles (row block I, column block J), one wave per tile: lane l
// owns rows I*SY_T + 2l, +1 and reads each column of the tile with one 16-B
// load per lane (a 1 KB burst per instruction); its two row sums stay in
// registers.  The column sums (over the tile's 128 rows) go through LDS
// SY_C columns at a time: the lane's two products per column are written
// column-major, 64 / SY_C lanes per column add SY_C entries each and xor
// shuffles join them, while the next SY_C columns' loads are in flight.
// Measured at n = 8192: 16-B loads and one wave per tile 6.1 TB/s, against
// 5.0 TB/s for 8-B loads with two waves per tile (whose loads alone top out
// at 5.5).  Row sums land in rowP[J][i], column sums in colP[I][j] (one writer
// each); symv_fold adds them per output in a fixed order and applies alpha /
// beta.  Diagonal tiles count the diagonal once, in the row sum.  Ragged
// edge tiles and a misaligned A (16-B loads need an even lda) take a
// per-element tested path.  The same tiles give DTRMV / DTPMV x := op(T) x
// with only the row sums ('N') or only the column sums ('T'): one read of the
// triangle, no copy of x (the fold writes x after every tile has read it).
constexpr int SY_T = 128;  // tile rows and columns
constexpr int SY_C = 8;    // columns per LDS pass

// Tile kinds: interior (16-B loads), diagonal inside the matrix (16-B loads
// of the whole tile, the unreferenced half -- still inside the caller's
// lda x n array -- replaced by zeros with a select, so garbage there never
// reaches the arithmetic), and edge (ragged rows or columns, or a misaligned
// A: every element tested before it is loaded).
enum { SY_INTERIOR = 0, SY_DIAG = 1, SY_EDGE = 2 };
// What a tile contributes: DSYMV / DSPMV both sums (the diagonal in the row
// sum only); op(T) x for a stored triangle T (DTRMV / DTPMV) the row sums
// ('N') or the column sums ('T'), the diagonal once (1 when unit).
enum { SY_MV_SYM = 0, SY_MV_N = 1, SY_MV_T = 2 };

// Packed storage (DSPMV, PACKED = true): column j of the triangle starts at
// j (j + 1) / 2 (upper) or j (2n - j - 1) / 2 (lower, rows from j), at any
// 8-byte offset, so a lane owns rows I*SY_T + l and + 64 (two 8-B loads per
// column, each instruction 512 contiguous bytes) instead of a 16-B pair.
// Inside a non-edge diagonal tile the unreferenced elements' addresses still
// fall inside the packed array (those of the neighbouring columns); they are
// loaded and replaced by zeros like the full-storage ones.
template <bool PACKED, bool LOWER>
__device__ __forceinline__ long sy_at(long i, long j, long lda, int n)
{
    if (!PACKED) return i + j * lda;
    return LOWER ? i + j * (2 * (long)n - j - 1) / 2 : i + j * (j + 1) / 2;
}

template <int KIND, bool LOWER, bool PACKED, bool UNIT = false>
__device__ __forceinline__ void symv_load(d2 (&a)[SY_C], const double *A, long lda, long ra,
                                          long rb, long j0, int n, bool diag)
{
#pragma unroll
    for (int u = 0; u < SY_C; ++u) {
        const long j = j0 + u;
        if (KIND == SY_INTERIOR) {
            if (PACKED) {
                a[u].x = __builtin_nontemporal_load(A + sy_at<true, LOWER>(ra, j, lda, n));
                a[u].y = __builtin_nontemporal_load(A + sy_at<true, LOWER>(rb, j, lda, n));
            } else {
                a[u] = __builtin_nontemporal_load((const d2 *)(A + ra + j * lda));
            }
        } else if (KIND == SY_DIAG) {
            d2 v;
            if (PACKED) {
                v.x = __builtin_nontemporal_load(A + sy_at<true, LOWER>(ra, j, lda, n));
                v.y = __builtin_nontemporal_load(A + sy_at<true, LOWER>(rb, j, lda, n));
            } else {
                v = __builtin_nontemporal_load((const d2 *)(A + ra + j * lda));
            }
            a[u].x = (LOWER ? ra >= j : ra <= j) ? (UNIT && ra == j ? 1.0 : v.x) : 0.0;
            a[u].y = (LOWER ? rb >= j : rb <= j) ? (UNIT && rb == j ? 1.0 : v.y) : 0.0;
        } else {
            const bool in = j < n;
            const bool s0 = in && ra < n && (!diag || (LOWER ? ra >= j : ra <= j));
            const bool s1 = in && rb < n && (!diag || (LOWER ? rb >= j : rb <= j));
            a[u].x = s0 ? (UNIT && ra == j ? 1.0 : A[sy_at<PACKED, LOWER>(ra, j, lda, n)]) : 0.0;
            a[u].y = s1 ? (UNIT && rb == j ? 1.0 : A[sy_at<PACKED, LOWER>(rb, j, lda, n)]) : 0.0;
        }
    }
}

template <int KIND, bool LOWER, bool PACKED, int MODE, bool UNIT>
__device__ __forceinline__ void symv_tile_body(int n, int I, int J, const double *A,
                                               long lda, d2 xi, const double *xs, double *rowP,
                                               double *colP, double (*pr)[65])
{
    constexpr bool EDGE = KIND == SY_EDGE;
    const int lane = threadIdx.x;
    // this lane's two rows
    const long ra = (long)I * SY_T + (PACKED ? lane : 2 * lane), rb = PACKED ? ra + 64 : ra + 1;
    const long j0 = (long)J * SY_T;
    const bool diag = I == J;
    double r0 = 0, r1 = 0;
    // one pass: row sums and column products of `cur` (columns c0 ..), the
    // next pass's columns loaded into `nxt`, then the column sums
    auto pass = [&](int c0, d2 (&cur)[SY_C], d2 (&nxt)[SY_C]) {
#pragma unroll
        for (int u = 0; u < SY_C; ++u) {
            if (MODE != SY_MV_T) {
                const double xj = xs[c0 + u];
                r0 = fma(cur[u].x, xj, r0);
                r1 = fma(cur[u].y, xj, r1);
            }
            if (MODE != SY_MV_N) {
                double px = cur[u].x * xi.x, py = cur[u].y * xi.y;
                if (MODE == SY_MV_SYM && KIND != SY_INTERIOR && diag) {
                    // the diagonal element: row sum only
                    const long j = j0 + c0 + u;
                    if (ra == j) px = 0.0;
                    if (rb == j) py = 0.0;
                }
                pr[u][lane] = px + py;
            }
        }
        // the next pass's loads fly while this pass's column sums are formed
        if (c0 + SY_C < SY_T)
            symv_load<KIND, LOWER, PACKED, UNIT>(nxt, A, lda, ra, rb, j0 + c0 + SY_C, n, diag);
        if constexpr (MODE == SY_MV_N) return;  // no column sums
        lds_barrier();
        {
            const int c = lane % SY_C, part = lane / SY_C;
            double sum = 0;
#pragma unroll
            for (int rr = 0; rr < SY_C; ++rr) sum += pr[c][part * SY_C + rr];
#pragma unroll
            for (int off = SY_C; off < 64; off *= 2) sum += __shfl_xor(sum, off);
            const long j = j0 + c0 + lane;
            if (lane < SY_C && (!EDGE || j < n)) colP[(long)I * n + j] = sum;
        }
        lds_barrier();  // pr is rewritten by the next pass
    };
    d2 a[SY_C], b[SY_C];
    symv_load<KIND, LOWER, PACKED, UNIT>(a, A, lda, ra, rb, j0, n, diag);
    lds_barrier();  // xs is in LDS; the first pass's loads stay in flight
    // The copy at the end of a pass waits for the next pass's loads, after
    // they have overlapped the column sums.  Measured at n = 8192: 49.9 us
    // per call; two passes per trip with the buffers' roles swapped (no
    // copies) 57.0 us, all passes unrolled 162 us.
#pragma unroll 1
    for (int c0 = 0; c0 < SY_T; c0 += SY_C) {
        pass(c0, a, b);
#pragma unroll
        for (int u = 0; u < SY_C; ++u) a[u] = b[u];
    }
    if (MODE == SY_MV_T) return;
    if (!EDGE || ra < n) rowP[(long)J * n + ra] = r0;
    if (!EDGE || rb < n) rowP[(long)J * n + rb] = r1;
}

template <bool LOWER, bool PACKED, int MODE, bool UNIT>  // template parameters: as run-time
                                    // values the triangle tests turn into branches that wait on every load
__global__ __launch_bounds__(64) void symv_tile(int n, int nb, int vec, const double *A,
                                                long lda, const double *x, long incx, long x0,
                                                double *rowP, double *colP)
{
    // stored tile t -> (p, q), q <= p; lower: (I, J) = (p, q), upper: (q, p).
    // The nb(nb-1)/2 off-diagonal tiles come first and the half-full
    // diagonal tiles last, so the final, partly empty round of workgroups is
    // made of the cheapest tiles.
    const long t = blockIdx.x, noff = (long)nb * (nb - 1) / 2;
    int p, q;
    if (t >= noff) {
        p = q = (int)(t - noff);
    } else {
        // strictly lower tiles in order (1,0), (2,0), (2,1), (3,0), ...
        p = (int)((sqrt(8.0 * (double)t + 1.0) + 1.0) * 0.5);
        while ((long)p * (p - 1) / 2 > t) --p;
        while ((long)(p + 1) * p / 2 <= t) ++p;
        q = (int)(t - (long)p * (p - 1) / 2);
    }
    const int I = LOWER ? p : q, J = LOWER ? q : p;
    const int lane = threadIdx.x;
    const long ra = (long)I * SY_T + (PACKED ? lane : 2 * lane), rb = PACKED ? ra + 64 : ra + 1;
    const long j0 = (long)J * SY_T;

    __shared__ double xs[SY_T];
    __shared__ double pr[SY_C][65];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const long j = j0 + lane + 64 * h;
        xs[lane + 64 * h] = j < n ? x[x0 + j * incx] : 0.0;
    }
    d2 xi;
    xi.x = ra < n ? x[x0 + ra * incx] : 0.0;
    xi.y = rb < n ? x[x0 + rb * incx] : 0.0;
    const bool edge = !vec || (long)(I + 1) * SY_T > n || (long)(J + 1) * SY_T > n;
    if (edge)
        symv_tile_body<SY_EDGE, LOWER, PACKED, MODE, UNIT>(n, I, J, A, lda, xi, xs, rowP, colP, pr);
    else if (I == J)
        symv_tile_body<SY_DIAG, LOWER, PACKED, MODE, UNIT>(n, I, J, A, lda, xi, xs, rowP, colP, pr);
    else
        symv_tile_body<SY_INTERIOR, LOWER, PACKED, MODE, UNIT>(n, I, J, A, lda, xi, xs, rowP, colP,
                                                              pr);
}

// Fold: 8 lanes per output.  Output k (block b) sums its row partials
// (column blocks beside b) followed by its column partials (row blocks
// beside b); lane g adds entries g, g+8, ... in that order and a fixed xor
// tree joins the 8 lanes, so the summation order depends only on n.
constexpr int SY_FL = 8;

__global__ __launch_bounds__(256) void symv_fold(int n, int lower, int nb, int mode,
                                                 const double *rowP, const double *colP,
                                                 double alpha, double beta, double *y, long incy,
                                                 long y0)
{
    const long gid = (long)blockIdx.x * 256 + threadIdx.x;
    const long k = gid / SY_FL;
    const int g = (int)(gid % SY_FL);
    double s = 0;
    if (k < n) {
        const int b = (int)(k / SY_T);
        const int j_lo = lower ? 0 : b, nrow = mode == SY_MV_T ? 0 : lower ? b + 1 : nb - b;
        const int i_lo = lower ? b : 0, ncol = mode == SY_MV_N ? 0 : lower ? nb - b : b + 1;
        const int tot = nrow + ncol;
        // 8 loads in flight per lane, then added in entry order
        for (int e0 = g; e0 < tot; e0 += 8 * SY_FL) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = e0 + u * SY_FL, ec = min(e, tot - 1);  // ec: always in bounds
                const double *src = ec < nrow ? rowP + (long)(j_lo + ec) * n
                                              : colP + (long)(i_lo + ec - nrow) * n;
                v[u] = e < tot ? src[k] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) s += v[u];
        }
    }
    s += __shfl_xor(s, 1);
    s += __shfl_xor(s, 2);
    s += __shfl_xor(s, 4);
    if (k < n && g == 0) {
        double *yp = y + y0 + k * incy;
        *yp = beta == 0.0 ? alpha * s : fma(beta, *yp, alpha * s);
    }
}

// DSYR: A += alpha*x*x' on the uplo triangle, HBM-bound (the triangle read
// and written once).  A workgroup owns a 512-row x 16-column block: lane t
// holds rows 2t, 2t+1 of the block, loads the 16 columns' pairs first (16-B
// loads) and then updates and stores them.  Blocks wholly inside the stored
// triangle and the matrix take that path unconditionally (a column with
// x(j) == 0 keeps its values, as netlib skips it); blocks cut by the diagonal
// or the matrix edge, or a misaligned A, test every element; blocks wholly
// outside the triangle return at once.
constexpr int SR_ROWS = 512;  // rows per workgroup (256 threads x 2)
constexpr int SR_COLS = 16;   // columns per workgroup

// TWO: the rank-2 update of DSYR2, A += x*(alpha y(j)) + y*(alpha x(j)) per
// column (netlib's order), a column skipped when x(j) and y(j) are both 0.
template <bool LOWER, bool TWO>
__global__ __launch_bounds__(256) void syr_kernel(int n, double alpha, const double *x, long incx,
                                                  long x0, const double *y, long incy, long y0,
                                                  double *A, long lda, int vec)
{
    const long r0 = (long)blockIdx.x * SR_ROWS, c0 = (long)blockIdx.y * SR_COLS;
    const long rl = r0 + SR_ROWS - 1, cl = c0 + SR_COLS - 1;
    if (LOWER ? rl < c0 : r0 > cl) return;  // no stored element in the block
    const long i = r0 + 2 * threadIdx.x;
    const bool full = vec && rl < n && cl < n && (LOWER ? r0 >= cl : rl <= c0);
    double xt[SR_COLS], yt[SR_COLS];
#pragma unroll
    for (int u = 0; u < SR_COLS; ++u) {
        xt[u] = c0 + u < n ? x[x0 + (c0 + u) * incx] : 0.0;
        yt[u] = TWO && c0 + u < n ? y[y0 + (c0 + u) * incy] : 0.0;
    }
    if (full) {
        const double xi0 = x[x0 + i * incx], xi1 = x[x0 + (i + 1) * incx];
        const double yi0 = TWO ? y[y0 + i * incy] : 0.0, yi1 = TWO ? y[y0 + (i + 1) * incy] : 0.0;
        d2 v[SR_COLS];
#pragma unroll
        for (int u = 0; u < SR_COLS; ++u)
            v[u] = __builtin_nontemporal_load((const d2 *)(A + i + (c0 + u) * lda));
#pragma unroll
        for (int u = 0; u < SR_COLS; ++u) {
            d2 w;
            if (TWO) {
                const double t1 = alpha * yt[u], t2 = alpha * xt[u];
                const bool on = xt[u] != 0.0 || yt[u] != 0.0;
                w.x = on ? fma(yi0, t2, fma(xi0, t1, v[u].x)) : v[u].x;
                w.y = on ? fma(yi1, t2, fma(xi1, t1, v[u].y)) : v[u].y;
            } else {
                const double t = alpha * xt[u];
                w.x = xt[u] != 0.0 ? fma(xi0, t, v[u].x) : v[u].x;
                w.y = xt[u] != 0.0 ? fma(xi1, t, v[u].y) : v[u].y;
            }
            __builtin_nontemporal_store(w, (d2 *)(A + i + (c0 + u) * lda));
        }
        return;
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const long r = i + h;
        if (r >= n) break;
        const double xr = x[x0 + r * incx], yr = TWO ? y[y0 + r * incy] : 0.0;
        for (int u = 0; u < SR_COLS; ++u) {
            const long j = c0 + u;
            if (j >= n || (LOWER ? r < j : r > j) || (xt[u] == 0.0 && yt[u] == 0.0)) continue;
            double *p = A + r + j * lda;
            *p = TWO ? fma(yr, alpha * xt[u], fma(xr, alpha * yt[u], *p))
                     : fma(xr, alpha * xt[u], *p);
        }
    }
}

// DSPR / DSPR2 on packed storage: column j of an upper triangle holds rows
// 0..j at AP[j(j+1)/2 ..], of a lower one rows j..n-1 at AP[j(2n-j+1)/2 ..];
// within a column consecutive rows are consecutive elements, so the 256
// threads of a workgroup (one row each, 16 columns) access memory coalesced.
template <bool LOWER, bool TWO>
__global__ __launch_bounds__(256) void spr_kernel(int n, double alpha, const double *x, long incx,
                                                  long x0, const double *y, long incy, long y0,
                                                  double *AP)
{
    const long r0 = (long)blockIdx.x * 256, c0 = (long)blockIdx.y * SR_COLS;
    if (LOWER ? r0 + 255 < c0 : r0 > c0 + SR_COLS - 1) return;
    const long r = r0 + threadIdx.x;
    if (r >= n) return;
    const double xr = x[x0 + r * incx], yr = TWO ? y[y0 + r * incy] : 0.0;
    for (int u = 0; u < SR_COLS; ++u) {
        const long j = c0 + u;
        if (j >= n || (LOWER ? r < j : r > j)) continue;
        const double xj = x[x0 + j * incx], yj = TWO ? y[y0 + j * incy] : 0.0;
        if (xj == 0.0 && yj == 0.0) continue;
        const long k = LOWER ? j * (2L * n - j + 1) / 2 + (r - j) : j * (j + 1) / 2 + r;
        AP[k] = TWO ? fma(yr, alpha * xj, fma(xr, alpha * yj, AP[k])) : fma(xr, alpha * xj, AP[k]);
    }
}

inline long start_of(int n, int inc) { return inc < 0 ? -(long)(n - 1) * inc : 0; }
inline bool isN(char c) { return c == 'N' || c == 'n'; }

// Column-panel plan of DGEMV 'N' (and DGER): nrb row blocks of GN_ROWS rows
// times ncb panels of cbw columns, about `target` workgroups (DGEMV: 512,
// two per CU; DGER: 256, one per CU, where its read-modify-write stream
// measured 2% faster), panels at least 16 columns wide.  With ncb > 1 the
// DGEMV panels' partial sums (ncb x m doubles) are folded in panel order by
// a second kernel.
struct GemvNPlan {
    int nrb, ncb, cbw;
};
GemvNPlan gemv_n_plan(int m, int n, int target = 512)
{
    GemvNPlan p;
    p.nrb = ceil_div(m, GN_ROWS);
    p.ncb = ceil_div(target, p.nrb);
    const int maxcb = ceil_div(n, 16);
    if (p.ncb > maxcb) p.ncb = maxcb;
    if (p.ncb < 1) p.ncb = 1;
    p.cbw = ceil_div(n, p.ncb);
    p.ncb = ceil_div(n, p.cbw);
    return p;
}

// ---- scratch of the allocation-free DGEMV entry point ------------------------
// The panel partials of gpublas_dev_dgemv_noalloc live in this fixed array
// instead of a pool block.  ncb > 1 needs nrb <= 511, and then
// ncb * m <= ceil(512 / nrb) * GN_ROWS * nrb <= 512 * (512 + 511) = 523,776
// doubles, so 2^19 doubles (4 MiB per GPU, held until the process ends) fit
// every shape.  Host code reaches the array only through the address that
// noalloc_scratch_setup() takes with hipGetSymbolAddress on each device; the
// array's own name is a host-side stand-in that kernels cannot use.
constexpr size_t kNoallocPartials = size_t(1) << 19;
__device__ __align__(256) double scratch[kNoallocPartials];
constexpr int kScratchDevices = 64;
double *g_scratch_ptr[kScratchDevices];

// One DGEMV on stream s.  fixed_partials = nullptr: the partials come from
// the stream's scratch (dev_scratch, which may allocate); otherwise they are
// fixed_partials (kNoallocPartials doubles, no allocation, no profiling
// events).
int gemv_run(hipStream_t s, char trans, int m, int n, double alpha, const double *d_A, int lda,
             const double *d_x, int incx, double beta, double *d_y, int incy,
             double *fixed_partials)
{
    if (m <= 0 || n <= 0) return 0;
    const bool prof = g_prof_on && !fixed_partials;
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
        const GemvNPlan pl = gemv_n_plan(m, n);
        double *partials = nullptr;
        Scratch scr;
        if (pl.ncb > 1) {
            const size_t need = (size_t)pl.ncb * m;
            if (fixed_partials) {
                if (need > kNoallocPartials) return (int)hipErrorInvalidValue;
                partials = fixed_partials;
            } else {
                scr = dev_scratch(need * sizeof(double), s);
                partials = (double *)scr.p;
                if (!partials) return (int)hipErrorOutOfMemory;
            }
        }
        const int vec = (((uintptr_t)d_A & 15) == 0) && (lda % 2 == 0);
        if (prof) prof_begin(s, "dgemv_n");
        hipLaunchKernelGGL(gemv_n_stream, dim3((unsigned)pl.nrb * pl.ncb), dim3(256), 0, s, m, n,
                           alpha, d_A, (long)lda, d_x, (long)incx, x0, beta, d_y, (long)incy, y0,
                           pl.nrb, pl.cbw, pl.ncb == 1, vec, partials);
        if (prof) prof_end(s);
        if (pl.ncb > 1) {
            if (prof) prof_begin(s, "dgemv_n_fold");
            hipLaunchKernelGGL(gemv_n_fold, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, s, m,
                               pl.ncb, partials, alpha, beta, d_y, (long)incy, y0);
            if (prof) prof_end(s);
        }
        return (int)hipGetLastError();
    }
    const int vec = (((uintptr_t)d_A & 15) == 0) && (lda % 2 == 0) && (((uintptr_t)d_x & 15) == 0);
    int blocks = ceil_div(n, 4);
    if (blocks > 2048) blocks = 2048;
    if (prof) prof_begin(s, "dgemv_t");
    hipLaunchKernelGGL(gemv_t_kernel, dim3((unsigned)blocks), dim3(256), 0, s, m, n, alpha, d_A,
                       (long)lda, d_x, (long)incx, x0, beta, d_y, (long)incy, y0, vec);
    if (prof) prof_end(s);
    return (int)hipGetLastError();
}

}  // namespace

// Takes the device address of `scratch` on the current device (called by the
// library's init with that device current).  Returns 0 or the HIP error.
int noalloc_scratch_setup(int device)
{
    if (device < 0 || device >= kScratchDevices) return (int)hipErrorInvalidDevice;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return e != hipSuccess ? (int)e : (int)hipErrorInvalidSymbol;
    __atomic_store_n(&g_scratch_ptr[device], (double *)ptr, __ATOMIC_RELEASE);
    return 0;
}

extern "C" {

int gpublas_dev_dgemv(void *stream, char trans, int m, int n, double alpha, const double *d_A,
                      int lda, const double *d_x, int incx, double beta, double *d_y, int incy)
{
    return gemv_run(pick_stream(stream), trans, m, n, alpha, d_A, lda, d_x, incx, beta, d_y, incy,
                    nullptr);
}

int gpublas_dev_dgemv_noalloc(void *stream, char trans, int m, int n, double alpha,
                              const double *d_A, int lda, const double *d_x, int incx, double beta,
                              double *d_y, int incy)
{
    hipDevice_t dev = -1;
    if (hipStreamGetDevice((hipStream_t)stream, &dev) != hipSuccess || dev < 0 ||
        dev >= kScratchDevices)
        return (int)hipErrorInvalidDevice;
    double *partials = __atomic_load_n(&g_scratch_ptr[dev], __ATOMIC_ACQUIRE);
    if (!partials) {
        // gpublas_dev_init() takes the address; if it has not run, or could
        // not take it, try once more here with the stream's device current
        int cur = -1;
        if (!dev_ready() || hipGetDevice(&cur) != hipSuccess || cur != dev)
            return (int)hipErrorNotInitialized;
        const int e = noalloc_scratch_setup(dev);
        if (e) return e;
        partials = __atomic_load_n(&g_scratch_ptr[dev], __ATOMIC_ACQUIRE);
    }
    return gemv_run((hipStream_t)stream, trans, m, n, alpha, d_A, lda, d_x, incx, beta, d_y, incy,
                    partials);
}

int gpublas_dev_dger(void *stream, int m, int n, double alpha, const double *d_x, int incx,
                     const double *d_y, int incy, double *d_A, int lda)
{
    hipStream_t s = pick_stream(stream);
    if (m <= 0 || n <= 0 || alpha == 0.0) return 0;
    const GemvNPlan pl = gemv_n_plan(m, n, 256);
    const int vec = (((uintptr_t)d_A & 15) == 0) && (lda % 2 == 0);
    PROF_BEGIN(s, "dger");
    hipLaunchKernelGGL(ger_stream, dim3((unsigned)pl.nrb * pl.ncb), dim3(256), 0, s, m, n, alpha,
                       d_x, (long)incx, start_of(m, incx), d_y, (long)incy, start_of(n, incy), d_A,
                       (long)lda, pl.nrb, pl.cbw, vec);
    PROF_END(s);
    return (int)hipGetLastError();
}

}  // extern "C"

namespace {
// tiles + fold of DSYMV / DSPMV (mode SY_MV_SYM) and of DTRMV / DTPMV
// (SY_MV_N / SY_MV_T, the result written over x through the fold: every tile
// has read x before the fold starts), full or packed storage
int symv_run(hipStream_t s, int lower, bool packed, int mode, bool unit, int vec, int n,
             double alpha, const double *A, int lda, const double *x, int incx, double beta,
             double *y, int incy, long y0)
{
    const int nb = ceil_div(n, SY_T);
    const long tiles = (long)nb * (nb + 1) / 2;
    Scratch scr = dev_scratch(2 * (size_t)nb * n * sizeof(double), s);
    double *part = (double *)scr.p;
    if (!part) return (int)hipErrorOutOfMemory;
    double *rowP = part, *colP = part + (size_t)nb * n;
    static const char *const names[2][3] = {{"dsymv", "dtrmv", "dtrmv"}, {"dspmv", "dtpmv", "dtpmv"}};
    PROF_BEGIN(s, names[packed][mode]);
#define GB_SYT(L, P, M, U)                                                                        \
    hipLaunchKernelGGL((symv_tile<L, P, M, U>), dim3((unsigned)tiles), dim3(64), 0, s, n, nb, vec, \
                       A, (long)lda, x, (long)incx, start_of(n, incx), rowP, colP)
#define GB_SYT_M(L, P)                                                                            \
    {                                                                                             \
        if (mode == SY_MV_SYM) GB_SYT(L, P, SY_MV_SYM, false);                                    \
        else if (mode == SY_MV_N) { if (unit) GB_SYT(L, P, SY_MV_N, true); else GB_SYT(L, P, SY_MV_N, false); } \
        else { if (unit) GB_SYT(L, P, SY_MV_T, true); else GB_SYT(L, P, SY_MV_T, false); }       \
    }
    if (packed) {
        if (lower) GB_SYT_M(true, true) else GB_SYT_M(false, true)
    } else {
        if (lower) GB_SYT_M(true, false) else GB_SYT_M(false, false)
    }
#undef GB_SYT_M
#undef GB_SYT
    PROF_END(s);
    static const char *const fnames[2][3] = {{"dsymv_fold", "dtrmv_fold", "dtrmv_fold"},
                                             {"dspmv_fold", "dtpmv_fold", "dtpmv_fold"}};
    PROF_BEGIN(s, fnames[packed][mode]);
    hipLaunchKernelGGL(symv_fold, dim3((unsigned)ceil_div((long)n * SY_FL, 256)), dim3(256), 0, s,
                       n, lower, nb, mode, (const double *)rowP, (const double *)colP, alpha, beta,
                       y, (long)incy, y0);
    PROF_END(s);
    return (int)hipGetLastError();
}
}  // namespace

extern "C" {

int gpublas_dev_dsymv(void *stream, char uplo, int n, double alpha, const double *d_A, int lda,
                      const double *d_x, int incx, double beta, double *d_y, int incy)
{
    hipStream_t s = pick_stream(stream);
    if (n <= 0) return 0;
    const long y0 = start_of(n, incy);
    if (alpha == 0.0) {
        if (beta == 1.0) return 0;
        int blocks = ceil_div(n, 256);
        if (blocks > 4096) blocks = 4096;
        hipLaunchKernelGGL(scal_vec_kernel, dim3(blocks), dim3(256), 0, s, n, beta, d_y,
                           (long)incy, y0);
        return (int)hipGetLastError();
    }
    const int lower = uplo == 'L' || uplo == 'l';
    const int vec = (((uintptr_t)d_A & 15) == 0) && (lda % 2 == 0);
    return symv_run(s, lower, false, SY_MV_SYM, false, vec, n, alpha, d_A, lda, d_x, incx, beta, d_y,
                    incy, y0);
}

// DSPMV: the DSYMV tile kernel on packed storage (one read of the packed
// triangle; replaces cublasDspmv, blas_level2/pmv.cc)
int gpublas_dev_dspmv(void *stream, char uplo, int n, double alpha, const double *d_AP,
                      const double *d_x, int incx, double beta, double *d_y, int incy)
{
    hipStream_t s = pick_stream(stream);
    if (n <= 0) return 0;
    if (alpha == 0.0) return gpublas_dev_dsymv(s, uplo, n, alpha, nullptr, n, d_x, incx, beta, d_y, incy);
    const int lower = uplo == 'L' || uplo == 'l';
    return symv_run(s, lower, true, SY_MV_SYM, false, 1, n, alpha, d_AP, 0, d_x, incx, beta, d_y,
                    incy, start_of(n, incy));
}

// DTRMV / DTPMV x := op(T) x: the tile kernel's row sums ('N') or column sums
// ('T') over the stored triangle, read once; the fold writes x (replaces
// cublasDtrmv / cublasDtpmv, blas_level2/trmv.cc, tpmv.cc)
int gpublas_dev_dtrmv(void *stream, char uplo, char trans, char diag, int n, const double *d_A,
                      int lda, double *d_x, int incx)
{
    hipStream_t s = pick_stream(stream);
    if (n <= 0) return 0;
    const int lower = uplo == 'L' || uplo == 'l';
    const int mode = trans == 'N' || trans == 'n' ? SY_MV_N : SY_MV_T;
    const bool unit = diag == 'U' || diag == 'u';
    const int vec = (((uintptr_t)d_A & 15) == 0) && (lda % 2 == 0);
    return symv_run(s, lower, false, mode, unit, vec, n, 1.0, d_A, lda, d_x, incx, 0.0, d_x, incx,
                    start_of(n, incx));
}

int gpublas_dev_dtpmv(void *stream, char uplo, char trans, char diag, int n, const double *d_AP,
                      double *d_x, int incx)
{
    hipStream_t s = pick_stream(stream);
    if (n <= 0) return 0;
    const int lower = uplo == 'L' || uplo == 'l';
    const int mode = trans == 'N' || trans == 'n' ? SY_MV_N : SY_MV_T;
    const bool unit = diag == 'U' || diag == 'u';
    return symv_run(s, lower, true, mode, unit, 1, n, 1.0, d_AP, 0, d_x, incx, 0.0, d_x, incx,
                    start_of(n, incx));
}

int gpublas_dev_dsyr(void *stream, char uplo, int n, double alpha, const double *d_x, int incx,
                     double *d_A, int lda)
{
    hipStream_t s = pick_stream(stream);
    if (n <= 0 || alpha == 0.0) return 0;
    const bool lower = uplo == 'L' || uplo == 'l';
    const int vec = (((uintptr_t)d_A & 15) == 0) && (lda % 2 == 0);
    dim3 grid((unsigned)ceil_div(n, SR_ROWS), (unsigned)ceil_div(n, SR_COLS));
    PROF_BEGIN(s, "dsyr");
    if (lower)
        hipLaunchKernelGGL((syr_kernel<true, false>), grid, dim3(256), 0, s, n, alpha, d_x,
                           (long)incx, start_of(n, incx), d_x, 1L, 0L, d_A, (long)lda, vec);
    else
        hipLaunchKernelGGL((syr_kernel<false, false>), grid, dim3(256), 0, s, n, alpha, d_x,
                           (long)incx, start_of(n, incx), d_x, 1L, 0L, d_A, (long)lda, vec);
    PROF_END(s);
    return (int)hipGetLastError();
}

int gpublas_dev_dsyr2(void *stream, char uplo, int n, double alpha, const double *d_x, int incx,
                      const double *d_y, int incy, double *d_A, int lda)
{
    hipStream_t s = pick_stream(stream);
    if (n <= 0 || alpha == 0.0) return 0;
    const bool lower = uplo == 'L' || uplo == 'l';
    const int vec = (((uintptr_t)d_A & 15) == 0) && (lda % 2 == 0);
    dim3 grid((unsigned)ceil_div(n, SR_ROWS), (unsigned)ceil_div(n, SR_COLS));
    PROF_BEGIN(s, "dsyr2");
    if (lower)
        hipLaunchKernelGGL((syr_kernel<true, true>), grid, dim3(256), 0, s, n, alpha, d_x,
                           (long)incx, start_of(n, incx), d_y, (long)incy, start_of(n, incy), d_A,
                           (long)lda, vec);
    else
        hipLaunchKernelGGL((syr_kernel<false, true>), grid, dim3(256), 0, s, n, alpha, d_x,
                           (long)incx, start_of(n, incx), d_y, (long)incy, start_of(n, incy), d_A,
                           (long)lda, vec);
    PROF_END(s);
    return (int)hipGetLastError();
}

static int spr_launch(hipStream_t s, bool two, char uplo, int n, double alpha, const double *d_x,
                      int incx, const double *d_y, int incy, double *d_AP)
{
    if (n <= 0 || alpha == 0.0) return 0;
    const bool lower = uplo == 'L' || uplo == 'l';
    dim3 grid((unsigned)ceil_div(n, 256), (unsigned)ceil_div(n, SR_COLS));
    const long x0 = start_of(n, incx), y0 = two ? start_of(n, incy) : 0;
    const long iy = two ? incy : 1;
    const double *y = two ? d_y : d_x;
    PROF_BEGIN(s, two ? "dspr2" : "dspr");
    if (lower && two)
        hipLaunchKernelGGL((spr_kernel<true, true>), grid, dim3(256), 0, s, n, alpha, d_x,
                           (long)incx, x0, y, iy, y0, d_AP);
    else if (lower)
        hipLaunchKernelGGL((spr_kernel<true, false>), grid, dim3(256), 0, s, n, alpha, d_x,
                           (long)incx, x0, y, iy, y0, d_AP);
    else if (two)
        hipLaunchKernelGGL((spr_kernel<false, true>), grid, dim3(256), 0, s, n, alpha, d_x,
                           (long)incx, x0, y, iy, y0, d_AP);
    else
        hipLaunchKernelGGL((spr_kernel<false, false>), grid, dim3(256), 0, s, n, alpha, d_x,
                           (long)incx, x0, y, iy, y0, d_AP);
    PROF_END(s);
    return (int)hipGetLastError();
}

int gpublas_dev_dspr(void *stream, char uplo, int n, double alpha, const double *d_x, int incx,
                     double *d_AP)
{
    return spr_launch(pick_stream(stream), false, uplo, n, alpha, d_x, incx, nullptr, 1, d_AP);
}

int gpublas_dev_dspr2(void *stream, char uplo, int n, double alpha, const double *d_x, int incx,
                      const double *d_y, int incy, double *d_AP)
{
    return spr_launch(pick_stream(stream), true, uplo, n, alpha, d_x, incx, d_y, incy, d_AP);
}

}  // extern "C"
