// ozaki.hip -- opt-in split-precision DGEMM / DSYRK on BF16 MFMA.
//
// The error-free splitting of Ozaki, Ogita, Oishi & Rump ("Error-free
// transformations of matrix multiplication by using fast routines of matrix
// multiplication and its applications", Numer. Algorithms 59, 2012), in the
// integer-slice form:
//  1. every row i of op(A) is scaled by 2^-ea_i and every column j of op(B)
//     by 2^-eb_j so its largest entry has magnitude below 1 (exact);
//  2. each scaled entry x is cut into s signed 8-bit integers:
//       t = 256 x;  q_p = trunc(t);  x = t - q_p      (p = 1..s, all exact)
//     so a = 2^ea sum_p q_p 2^-8p + r with |r| < 2^(ea-8s).  |q_p| <= 255 is
//     exactly representable in BF16 (8 significant bits);
//  3. for every slice pair with p + q <= s + 1 the integer product
//     Q_A^p Q_B^q is formed on v_mfma_f32_32x32x16_bf16.  Over a k-chunk of
//     256 every partial sum is an integer of magnitude <= 255^2 * 256 < 2^24,
//     so the FP32 accumulation is exact;
//  4. after each chunk and pair the FP32 tile is added, scaled by
//     2^-8(p+q), into an FP64 accumulator; the epilogue applies
//     2^(ea_i + eb_j), alpha and beta in FP64.
// The only approximations are the dropped pairs (p + q > s + 1), the slice
// remainders and the FP64 additions.  FP64 is used exactly where the scheme
// needs it: the scaling/splitting pass and the final summation.
#include "devcommon.hpp"
#include "gpublas_dev.h"
#include <math.h>

namespace {

typedef __bf16 bf16;
typedef short v8s __attribute__((ext_vector_type(8)));  // 16 B of bf16 for loads
constexpr int OZ_BM = 128, OZ_BN = 128, OZ_BKT = 32, OZ_CHUNK = 256, OZ_MAXS = 8;
constexpr int OZ_LDSK = OZ_BKT + 8;  // bf16 per LDS row (16-B pad)

struct OzArgs {
    int M, N, Kp;           // Kp: K rounded up to a multiple of OZ_BKT (zero slices)
    int s, npairs;
    const bf16 *SA, *SB;    // [s][M][Kp], [s][N][Kp], k contiguous
    const int *ea, *eb;     // row exponents of op(A), column exponents of op(B)
    double alpha, beta;
    double *C;
    long ldc;
    int tiles_m;
    int tri;                // 0 rectangle, 1 lower, 2 upper (SYRK)
    unsigned char pp[64], pq[64];  // slice pairs
};

// ---- exponents -----------------------------------------------------------
// e such that max|x| < 2^e (0 for an all-zero line); a non-finite entry sets
// *bad so the caller can fall back to plain FP64.
__device__ __forceinline__ int exp_of(double mx)
{
    if (mx == 0.0) return 0;
    int e;
    frexp(mx, &e);  // mx = f * 2^e, f in [0.5, 1)
    return e;
}

// line r of an operand stored so that consecutive r are contiguous
// (element (r, k) at X[r + k*ld]): one thread per line, coalesced over r.
__global__ void exp_rows_strided(int R, int K, const double *X, long ld, int *e, int *bad)
{
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    double mx = 0;
    for (int k = 0; k < K; ++k) {
        const double v = fabs(X[r + (long)k * ld]);
        if (!(v <= 1.7976931348623157e308)) *bad = 1;
        mx = fmax(mx, v);
    }
    e[r] = exp_of(mx);
}

// line r stored contiguously (element (r, k) at X[k + r*ld]): one wave per line.
__global__ void exp_rows_contig(int R, int K, const double *X, long ld, int *e, int *bad)
{
    const int lane = threadIdx.x & 63;
    const long r = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= R) return;
    double mx = 0;
    for (int k = lane; k < K; k += 64) {
        const double v = fabs(X[k + r * ld]);
        if (!(v <= 1.7976931348623157e308)) *bad = 1;
        mx = fmax(mx, v);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if (lane == 0) e[r] = exp_of(mx);
}

// ---- splitting -----------------------------------------------------------
// 32 (r) x 64 (k) tile through LDS so both the FP64 read (either layout) and
// the BF16 slice writes (k contiguous) are coalesced.
__global__ __launch_bounds__(256) void split_kernel(int R, int K, int Kp, int s, const double *X,
                                                    long ld, int rcontig, const int *e, bf16 *S)
{
    __shared__ double t[32][65];
    const int r0 = blockIdx.x * 32, k0 = blockIdx.y * 64;
    for (int idx = threadIdx.x; idx < 32 * 64; idx += 256) {
        int rr, kk;
        if (rcontig) { rr = idx % 32; kk = idx / 32; }  // r fastest: X[r + k*ld]
        else { kk = idx % 64; rr = idx / 64; }          // k fastest: X[k + r*ld]
        const long r = r0 + rr, k = k0 + kk;
        double v = 0;
        if (r < R && k < K) v = rcontig ? X[r + k * ld] : X[k + r * ld];
        t[rr][kk] = v;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < 32 * 32; idx += 256) {
        const int kk = (idx % 32) * 2, rr = idx / 32;
        const long r = r0 + rr, k = k0 + kk;
        if (r >= R || k >= Kp) continue;
        const int er = e[r];
        double x0 = ldexp(t[rr][kk], -er), x1 = ldexp(t[rr][kk + 1], -er);
        for (int p = 0; p < s; ++p) {
            const double u0 = x0 * 256.0, u1 = x1 * 256.0;
            const double q0 = trunc(u0), q1 = trunc(u1);
            x0 = u0 - q0;
            x1 = u1 - q1;
            bf16 *o = S + ((long)p * R + r) * Kp + k;
            o[0] = (bf16)(float)q0;
            o[1] = (bf16)(float)q1;
        }
    }
}

// ---- slice products ----------------------------------------------------------
__device__ __forceinline__ void tri_decode(int b, int &hi, int &lo)
{
    int h = (int)((sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
    while (h * (h + 1) / 2 > b) --h;
    while ((h + 1) * (h + 2) / 2 <= b) ++h;
    hi = h;
    lo = b - h * (h + 1) / 2;
}

__global__ __launch_bounds__(256) void ozaki_gemm_kernel(OzArgs g)
{
    __shared__ __attribute__((aligned(16))) bf16 As[2][OZ_BM][OZ_LDSK];
    __shared__ __attribute__((aligned(16))) bf16 Bs[2][OZ_BN][OZ_LDSK];
    int tm, tn;
    if (g.tri == 0) {
        tm = blockIdx.x % g.tiles_m;
        tn = blockIdx.x / g.tiles_m;
    } else {
        int hi, lo;
        tri_decode(blockIdx.x, hi, lo);
        if (g.tri == 1) { tm = hi; tn = lo; } else { tm = lo; tn = hi; }
    }
    const int m0 = tm * OZ_BM, n0 = tn * OZ_BN;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    v16f acc32[2][2];
    double acc64[2][2][16];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            acc32[i][j] = v16f{};
#pragma unroll
            for (int r = 0; r < 16; ++r) acc64[i][j][r] = 0.0;
        }

    const int nkt_total = g.Kp / OZ_BKT;               // k-tiles over all of K
    const int kt_per_chunk = OZ_CHUNK / OZ_BKT;
    const int nchunks = (nkt_total + kt_per_chunk - 1) / kt_per_chunk;
    // flattened step list: chunk-major, then pair, then k-tile within the chunk
    auto chunk_tiles = [&](int c) { return min(kt_per_chunk, nkt_total - c * kt_per_chunk); };

    v8s ra[2], rb[2];
    auto load = [&](int c, int pr, int kt) {
        const int k = (c * kt_per_chunk + kt) * OZ_BKT;
        const bf16 *sa = g.SA + (long)g.pp[pr] * g.M * g.Kp;
        const bf16 *sb = g.SB + (long)g.pq[pr] * g.N * g.Kp;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int v = threadIdx.x + r * 256;
            const int row = v >> 2, part = v & 3;
            const long ga = m0 + row, gb = n0 + row;
            ra[r] = ga < g.M ? *(const v8s *)(sa + ga * g.Kp + k + part * 8) : v8s{};
            rb[r] = gb < g.N ? *(const v8s *)(sb + gb * g.Kp + k + part * 8) : v8s{};
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int v = threadIdx.x + r * 256;
            const int row = v >> 2, part = v & 3;
            *(v8s *)&As[buf][row][part * 8] = ra[r];
            *(v8s *)&Bs[buf][row][part * 8] = rb[r];
        }
    };

    int c = 0, pr = 0, kt = 0, buf = 0;
    load(0, 0, 0);
    store(0);
    __syncthreads();
    while (true) {
        // next step
        int c1 = c, pr1 = pr, kt1 = kt + 1;
        if (kt1 == chunk_tiles(c1)) {
            kt1 = 0;
            if (++pr1 == g.npairs) {
                pr1 = 0;
                ++c1;
            }
        }
        const bool more = c1 < nchunks;
        if (more) load(c1, pr1, kt1);
#pragma unroll
        for (int kk = 0; kk < OZ_BKT; kk += 16) {
            v8bf fa[2], fb[2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                fa[i] = *(const v8bf *)&As[buf][wm * 64 + i * 32 + (lane & 31)][kk + 8 * (lane >> 5)];
#pragma unroll
            for (int j = 0; j < 2; ++j)
                fb[j] = *(const v8bf *)&Bs[buf][wn * 64 + j * 32 + (lane & 31)][kk + 8 * (lane >> 5)];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc32[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fb[j], fa[i], acc32[i][j],
                                                                          0, 0, 0);
        }
        if (kt + 1 == chunk_tiles(c)) {
            // end of this (chunk, pair): fold the exact FP32 integers into FP64
            const double sc = ldexp(1.0, -8 * (g.pp[pr] + g.pq[pr] + 2));
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
#pragma unroll
                    for (int r = 0; r < 16; ++r) acc64[i][j][r] = fma((double)acc32[i][j][r], sc, acc64[i][j][r]);
                    acc32[i][j] = v16f{};
                }
        }
        if (!more) break;
        store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
        c = c1;
        pr = pr1;
        kt = kt1;
    }

    // epilogue: element r of tile (i, j) is C[row][col] with
    //   row = m0 + wm*64 + i*32 + lane%32, col = n0 + wn*64 + j*32 + 8*(r/4) + 4*(lane/32) + r%4
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const long gi = m0 + wm * 64 + i * 32 + (lane & 31);
        if (gi >= g.M) continue;
        const int ei = g.ea[gi];
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const long gj = n0 + wn * 64 + j * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (gj >= g.N) continue;
                if (g.tri == 1 && gi < gj) continue;
                if (g.tri == 2 && gi > gj) continue;
                double *cp = g.C + gi + gj * g.ldc;
                const double v = g.alpha * ldexp(acc64[i][j][r], ei + g.eb[gj]);
                *cp = g.beta == 0.0 ? v : fma(g.beta, *cp, v);
            }
    }
}

inline bool isN(char c) { return c == 'N' || c == 'n'; }

struct Split {
    bf16 *S;
    int *e;
};

// Exponents and slices of an R x K operand given by (X, ld, rcontig).
int make_split(hipStream_t st, int R, int K, int Kp, int s, const double *X, long ld, int rcontig,
               Split &out, int *bad)
{
    out.e = (int *)gpublas_dev_alloc((size_t)R * sizeof(int));
    out.S = (bf16 *)gpublas_dev_alloc((size_t)s * R * Kp * sizeof(bf16));
    if (!out.e || !out.S) return (int)hipErrorOutOfMemory;
    PROF_BEGIN(st, "ozaki_split");
    if (rcontig)
        hipLaunchKernelGGL(exp_rows_strided, dim3((unsigned)ceil_div(R, 256)), dim3(256), 0, st, R,
                           K, X, ld, out.e, bad);
    else
        hipLaunchKernelGGL(exp_rows_contig, dim3((unsigned)ceil_div(R, 4)), dim3(256), 0, st, R, K,
                           X, ld, out.e, bad);
    hipLaunchKernelGGL(split_kernel, dim3((unsigned)ceil_div(R, 32), (unsigned)ceil_div(Kp, 64)),
                       dim3(256), 0, st, R, K, Kp, s, X, ld, rcontig, out.e, out.S);
    PROF_END(st);
    return (int)hipGetLastError();
}

int run_products(hipStream_t st, int M, int N, int Kp, int s, const Split &a, const Split &b,
                 double alpha, double beta, double *C, int ldc, int tri)
{
    OzArgs g;
    g.M = M;
    g.N = N;
    g.Kp = Kp;
    g.s = s;
    g.SA = a.S;
    g.SB = b.S;
    g.ea = a.e;
    g.eb = b.e;
    g.alpha = alpha;
    g.beta = beta;
    g.C = C;
    g.ldc = ldc;
    g.tri = tri;
    g.npairs = 0;
    for (int p = 0; p < s; ++p)
        for (int q = 0; q < s; ++q)
            if (p + q <= s - 1) {  // 1-based p' + q' <= s + 1
                g.pp[g.npairs] = (unsigned char)p;
                g.pq[g.npairs] = (unsigned char)q;
                ++g.npairs;
            }
    g.tiles_m = ceil_div(M, OZ_BM);
    const long tn = ceil_div(N, OZ_BN);
    const long blocks = tri ? (long)g.tiles_m * (g.tiles_m + 1) / 2 : (long)g.tiles_m * tn;
    PROF_BEGIN(st, "ozaki_gemm");
    hipLaunchKernelGGL(ozaki_gemm_kernel, dim3((unsigned)blocks), dim3(256), 0, st, g);
    PROF_END(st);
    return (int)hipGetLastError();
}

}  // namespace

int gpublas_gemm_d(hipStream_t s, char ta, char tb, int m, int n, int k, double alpha,
                   const double *A, int lda, const double *B, int ldb, double beta, double *C,
                   int ldc);
int gpublas_dev_dsyrk(void *stream, char uplo, char trans, int n, int k, double alpha,
                      const double *d_A, int lda, double beta, double *d_C, int ldc);

extern "C" {

int gpublas_dev_dgemm_split(void *stream, int slices, char transa, char transb, int m, int n,
                            int k, double alpha, const double *d_A, int lda, const double *d_B,
                            int ldb, double beta, double *d_C, int ldc)
{
    hipStream_t st = pick_stream(stream);
    if (m <= 0 || n <= 0) return 0;
    if (slices < 1 || slices > OZ_MAXS) return (int)hipErrorInvalidValue;
    if (alpha == 0.0 || k <= 0)
        return gpublas_gemm_d(st, transa, transb, m, n, k, alpha, d_A, lda, d_B, ldb, beta, d_C, ldc);
    const int Kp = ceil_div(k, OZ_BKT) * OZ_BKT;
    int *bad = (int *)dev_workspace(sizeof(int));
    (void)hipMemsetAsync(bad, 0, sizeof(int), st);
    Split a{}, b{};
    // op(A) rows: A 'N' -> X[r + k*lda] (rows contiguous); 'T' -> X[k + r*lda]
    int err = make_split(st, m, k, Kp, slices, d_A, lda, isN(transa), a, bad);
    // op(B) columns: B 'N' -> element (k, j) at B[k + j*ldb] (k contiguous); 'T' -> B[j + k*ldb]
    if (!err) err = make_split(st, n, k, Kp, slices, d_B, ldb, !isN(transb), b, bad);
    int hbad = 0;
    if (!err) err = (int)hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st);
    if (!err) err = (int)hipStreamSynchronize(st);
    if (!err) {
        if (hbad)  // Inf / NaN input: the splitting is undefined, use the FP64 kernel
            err = gpublas_gemm_d(st, transa, transb, m, n, k, alpha, d_A, lda, d_B, ldb, beta, d_C, ldc);
        else
            err = run_products(st, m, n, Kp, slices, a, b, alpha, beta, d_C, ldc, 0);
    }
    (void)hipStreamSynchronize(st);
    gpublas_dev_release(a.S);
    gpublas_dev_release(a.e);
    gpublas_dev_release(b.S);
    gpublas_dev_release(b.e);
    return err;
}

int gpublas_dev_dsyrk_split(void *stream, int slices, char uplo, char trans, int n, int k,
                            double alpha, const double *d_A, int lda, double beta, double *d_C,
                            int ldc)
{
    hipStream_t st = pick_stream(stream);
    if (n <= 0) return 0;
    if (slices < 1 || slices > OZ_MAXS) return (int)hipErrorInvalidValue;
    if (alpha == 0.0 || k <= 0)
        return gpublas_dev_dsyrk(st, uplo, trans, n, k, alpha, d_A, lda, beta, d_C, ldc);
    const int Kp = ceil_div(k, OZ_BKT) * OZ_BKT;
    int *bad = (int *)dev_workspace(sizeof(int));
    (void)hipMemsetAsync(bad, 0, sizeof(int), st);
    Split a{};
    // op(A) = A (trans N: rows contiguous) or A^T; op(B) = op(A)^T has the same lines
    int err = make_split(st, n, k, Kp, slices, d_A, lda, isN(trans), a, bad);
    int hbad = 0;
    if (!err) err = (int)hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st);
    if (!err) err = (int)hipStreamSynchronize(st);
    if (!err) {
        if (hbad) err = gpublas_dev_dsyrk(st, uplo, trans, n, k, alpha, d_A, lda, beta, d_C, ldc);
        else err = run_products(st, n, n, Kp, slices, a, a, alpha, beta, d_C, ldc,
                                (uplo == 'U' || uplo == 'u') ? 2 : 1);
    }
    (void)hipStreamSynchronize(st);
    gpublas_dev_release(a.S);
    gpublas_dev_release(a.e);
    return err;
}

}  // extern "C"
