// blas1.hip -- Level-1 streaming kernels on gfx950: DAXPY, DDOT, DNRM2,
// IDAMAX, and DSCAL, DCOPY, DSWAP, DROT, DASUM.
//
// Replaces the vendor calls the reference's (unbuilt) CBLAS wrappers make:
// cublasDaxpy (blas_level1/axpy.cc:49), cublasDdot (dot.cc:46), cublasDnrm2
// (nrm2.cc:39), cublasIdamax (amax.cc:40).
//
// The reductions are single launches.  Vectors of up to 65536 elements are
// reduced by one 512-thread workgroup (no partials, no tickets).
// Longer ones: every workgroup reduces its grid-stride share in a fixed
// order (per-lane sequential, then a wave64 shuffle tree, then the 4 waves
// in order), writes one partial, and the workgroup that draws the last
// ticket folds the partials in index order.  Either way the scalar goes
// straight into pinned host memory.  The grid depends only
// on n, so the result is bitwise reproducible run to run and independent of
// what else is on the GPU.  Tickets reset themselves, so no memset is needed
// between calls.
#include "devcommon.hpp"
#include "gpublas_dev.h"
#include <math.h>
#include <mutex>
#include <new>
#include <time.h>
#include <stdlib.h>
#include <string.h>

namespace {

constexpr int kSlots = 64;              // reduction slots per host thread
constexpr int kMaxBlocks = 1024;        // partials per slot
constexpr size_t kSlotBytes = kMaxBlocks * 16;
thread_local unsigned t_slot = 0;

__global__ void zero_ticket(unsigned *t)
{
    if (threadIdx.x == 0) *t = 0u;
}

// Partials and ticket of one reduction on stream st.  On the calling thread's
// own stream: one of its kSlots rotating slots (tickets kept zero between
// launches, calls ordered by the stream).  On another stream: a pool block of
// its own, whose ticket a one-thread kernel clears first (a fill kernel rather
// than a memset, so the call can be captured into a graph).
struct RedSlot {
    Scratch scr;
    void *partials = nullptr;
    unsigned *ticket = nullptr;
};

RedSlot next_slot(hipStream_t st)
{
    DevCtx *c = dev_ctx();
    RedSlot r;
    if (st == c->stream) {
        r.scr = dev_scratch(kSlots * kSlotBytes, st);
        if (!r.scr.p) return r;
        unsigned s = t_slot++ % kSlots;
        r.partials = (char *)r.scr.p + s * kSlotBytes;
        r.ticket = c->tickets + s;
        return r;
    }
    r.scr = dev_scratch(kSlotBytes + 256, st);
    if (!r.scr.p) return r;
    r.partials = r.scr.p;
    r.ticket = (unsigned *)((char *)r.scr.p + kSlotBytes);
    hipLaunchKernelGGL(zero_ticket, dim3(1), dim3(64), 0, st, r.ticket);
    return r;
}

int red_blocks(int n)
{
    long b = ((long)n + 2047) / 2048;  // 8 elements per thread
    if (b < 1) b = 1;
    if (b > kMaxBlocks) b = kMaxBlocks;
    return (int)b;
}

__device__ __forceinline__ bool last_block(unsigned *ticket)
{
    __shared__ bool is_last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        is_last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (is_last) __threadfence();
    return is_last;
}

__device__ __forceinline__ double block_sum(double v, double *red)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double r = red[0] + red[1] + red[2] + red[3];
    __syncthreads();
    return r;
}

// Scalar results land in pinned host memory; when the destination is the
// calling thread's result slot, the kernel also publishes a sequence number
// next to it (system-scope fence, then the flag) so the host can poll for
// completion instead of synchronising the stream.
__device__ __forceinline__ void publish(uint64_t *flag, uint64_t seq)
{
    if (!flag) return;
    __threadfence_system();
    __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

struct Pub {
    uint64_t *flag;
    uint64_t seq;
};

Pub pub_for(const void *h_result)
{
    DevCtx *c = dev_ctx();
    if (h_result == c->pinned) return Pub{(uint64_t *)((char *)c->pinned + 8), c->scalar_seq};
    return Pub{nullptr, 0};
}

// ---- DAXPY ------------------------------------------------------------------
__global__ void axpy_unit_v2(long n2, double a, const double2 *__restrict__ x, double2 *y)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long stride = (long)gridDim.x * blockDim.x;
    for (; i < n2; i += stride) {
        double2 xv = x[i], yv = y[i];
        yv.x = fma(a, xv.x, yv.x);
        yv.y = fma(a, xv.y, yv.y);
        y[i] = yv;
    }
}

__global__ void axpy_strided(int n, double a, const double *x, long incx, long x0, double *y,
                             long incy, long y0)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) y[y0 + i * incy] = fma(a, x[x0 + i * incx], y[y0 + i * incy]);
}

// ---- DDOT -------------------------------------------------------------------
__global__ __launch_bounds__(256) void dot_kernel(int n, const double *x, long incx, long x0,
                                                  const double *y, long incy, long y0,
                                                  double *partials, unsigned *ticket,
                                                  double *result, uint64_t *flag, uint64_t seq)
{
    __shared__ double red[4];
    double s = 0;
    long i = (long)blockIdx.x * 256 + threadIdx.x;
    const long stride = (long)gridDim.x * 256;
    if (incx == 1 && incy == 1 && x0 == 0 && y0 == 0) {
        for (; i < n; i += stride) s = fma(x[i], y[i], s);
    } else {
        for (; i < n; i += stride) s = fma(x[x0 + i * incx], y[y0 + i * incy], s);
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
    if (!last_block(ticket)) return;
    double t = 0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += 256) t += partials[b];
    t = block_sum(t, red);
    if (threadIdx.x == 0) {
        *result = t;
        publish(flag, seq);
        __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---- DNRM2: Blue's three-accumulator sum of squares ---------------------------
// Thresholds and scalings for IEEE double (Anderson, "Algorithm 978: Safe
// Scaling in the Level 1 BLAS", ACM TOMS 44(1), 2017): values above tbig are
// summed scaled down by sbig, values below tsml scaled up by ssml, the rest
// unscaled.  The three sums are associative, so they reduce like a dot.
__constant__ double kTsml = 0x1p-511, kTbig = 0x1p486, kSsml = 0x1p537, kSbig = 0x1p-538;

struct Acc3 {
    double sml, med, big;
};

__device__ __forceinline__ void acc3_add(Acc3 &a, double v)
{
    const double ax = fabs(v);
    if (ax > kTbig) {
        const double t = ax * kSbig;
        a.big = fma(t, t, a.big);
    } else if (ax < kTsml) {
        const double t = ax * kSsml;
        a.sml = fma(t, t, a.sml);
    } else {
        a.med = fma(ax, ax, a.med);
    }
}

__device__ double nrm2_finish(double asml, double amed, double abig)
{
    double scl, sumsq;
    if (abig > 0) {
        if (amed > 0 || isnan(amed)) abig += (amed * kSbig) * kSbig;
        scl = 1.0 / kSbig;
        sumsq = abig;
    } else if (asml > 0) {
        if (amed > 0 || isnan(amed)) {
            const double ymed = sqrt(amed), ysml = sqrt(asml) / kSsml;
            const double ymin = ysml > ymed ? ymed : ysml, ymax = ysml > ymed ? ysml : ymed;
            const double q = ymin / ymax;
            scl = 1.0;
            sumsq = ymax * ymax * (1.0 + q * q);
        } else {
            scl = 1.0 / kSsml;
            sumsq = asml;
        }
    } else {
        scl = 1.0;
        sumsq = amed;
    }
    return scl * sqrt(sumsq);
}

__global__ __launch_bounds__(256) void nrm2_kernel(int n, const double *x, long incx, long x0,
                                                   double *partials, unsigned *ticket,
                                                   double *result, uint64_t *flag, uint64_t seq)
{
    __shared__ double red[4];
    Acc3 a{0, 0, 0};
    long i = (long)blockIdx.x * 256 + threadIdx.x;
    const long stride = (long)gridDim.x * 256;
    for (; i < n; i += stride) acc3_add(a, x[x0 + i * incx]);
    a.sml = block_sum(a.sml, red);
    a.med = block_sum(a.med, red);
    a.big = block_sum(a.big, red);
    if (threadIdx.x == 0) {
        partials[3 * blockIdx.x + 0] = a.sml;
        partials[3 * blockIdx.x + 1] = a.med;
        partials[3 * blockIdx.x + 2] = a.big;
    }
    if (!last_block(ticket)) return;
    Acc3 t{0, 0, 0};
    for (int b = threadIdx.x; b < (int)gridDim.x; b += 256) {
        t.sml += partials[3 * b + 0];
        t.med += partials[3 * b + 1];
        t.big += partials[3 * b + 2];
    }
    t.sml = block_sum(t.sml, red);
    t.med = block_sum(t.med, red);
    t.big = block_sum(t.big, red);
    if (threadIdx.x == 0) {
        *result = nrm2_finish(t.sml, t.med, t.big);
        publish(flag, seq);
        __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---- IDAMAX -------------------------------------------------------------------
// Candidate (value, index): larger |x| wins, equal |x| keeps the smaller
// index, NaN is mapped below every number (netlib's strict '>' never moves to
// a NaN).  If x(1) itself is NaN netlib returns 1; the final block checks it.
struct Cand {
    double v;
    long i;
};

// IDAMAX (MIN = false) and IDAMIN (MIN = true): the candidate with the
// larger (smaller) |x_i|, ties to the smaller index.  A NaN element enters as
// -1 for the maximum and +inf for the minimum, so it is never preferred; the
// netlib loop's NaN-at-position-1 rule is applied at the end.
template <bool MIN>
__device__ __forceinline__ Cand better(Cand a, Cand b)
{
    if ((MIN ? b.v < a.v : b.v > a.v) || (b.v == a.v && b.i < a.i)) return b;
    return a;
}

template <bool MIN>
__device__ __forceinline__ Cand cand_of(long i, double v)
{
    double a = fabs(v);
    if (isnan(a)) a = MIN ? __builtin_inf() : -1.0;
    return Cand{a, i};
}

template <bool MIN>
__device__ __forceinline__ Cand cand_none()
{
    return Cand{MIN ? __builtin_inf() : -1.0, 0x7fffffffffffffffL};
}

template <bool MIN>
__device__ __forceinline__ long cand_result(Cand t, const double *x)
{
    long r = t.i + 1;
    if (isnan(x[0]) || (!MIN && t.v < 0)) r = 1;
    return r;
}

template <bool MIN>
__device__ __forceinline__ Cand wave_best(Cand c)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        Cand d;
        d.v = __shfl_xor(c.v, o, 64);
        d.i = __shfl_xor(c.i, o, 64);
        c = better<MIN>(c, d);
    }
    return c;
}

template <bool MIN>
__device__ Cand block_best(Cand c, Cand *red)
{
    c = wave_best<MIN>(c);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = c;
    __syncthreads();
    Cand r = red[0];
    r = better<MIN>(r, red[1]);
    r = better<MIN>(r, red[2]);
    r = better<MIN>(r, red[3]);
    __syncthreads();
    return r;
}

template <bool MIN>
__global__ __launch_bounds__(256) void iamax_kernel(int n, const double *x, long incx,
                                                    double *partials, unsigned *ticket,
                                                    int *result, uint64_t *flag, uint64_t seq)
{
    __shared__ Cand red[4];
    Cand c = cand_none<MIN>();
    long i = (long)blockIdx.x * 256 + threadIdx.x;
    const long stride = (long)gridDim.x * 256;
    for (; i < n; i += stride) c = better<MIN>(c, cand_of<MIN>(i, x[i * incx]));
    c = block_best<MIN>(c, red);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = c.v;
        ((long *)partials)[2 * blockIdx.x + 1] = c.i;
    }
    if (!last_block(ticket)) return;
    Cand t = cand_none<MIN>();
    for (int b = threadIdx.x; b < (int)gridDim.x; b += 256)
        t = better<MIN>(t, Cand{partials[2 * b], ((long *)partials)[2 * b + 1]});
    t = block_best<MIN>(t, red);
    if (threadIdx.x == 0) {
        *result = (int)cand_result<MIN>(t, x);
        publish(flag, seq);
        __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---- single-workgroup reductions for short vectors ---------------------------
// Below kSmallN one 512-thread workgroup reduces the whole vector: thread t
// takes elements 2t, 2t+1, 2t+1024, ... (16-B loads, eight per operand in
// flight), then a wave64 shuffle tree and the 8 waves in order.  No
// tickets, no partials: one memory round trip per 8192 elements and the
// result publication.  512 threads at <= 128 VGPRs take half of one CU's
// registers, so the resident server (the same code) leaves room on its CU
// for other kernels' waves.
constexpr int kSmallN = 1 << 16;
constexpr int kWgThreads = 512;              // threads of a single-workgroup reduction
constexpr int kWgWaves = kWgThreads / 64;
constexpr int kWgStride = 2 * kWgThreads;    // elements per sweep of the workgroup
constexpr int kWgU = 8;                      // 16-B loads in flight per operand
typedef double d2v __attribute__((ext_vector_type(2)));

template <typename F>
__device__ __forceinline__ void for_pairs(int n, F &&f)
{
    // calls f(i) for the element indices of this thread in increasing order
    int i = 2 * (int)threadIdx.x;
    for (; i < n; i += kWgStride) {
        f(i);
        if (i + 1 < n) f(i + 1);
    }
}

__device__ __forceinline__ double block_sum_wg(double v, double *red)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double r = 0;
#pragma unroll
    for (int w = 0; w < kWgWaves; ++w) r += red[w];
    __syncthreads();
    return r;
}

__device__ __forceinline__ void dot_wg(int n, const double *x, long incx, long x0, const double *y,
                                       long incy, long y0, int vec, double *result, uint64_t *flag,
                                       uint64_t seq)
{
    __shared__ double red[kWgWaves];
    double s = 0;
    if (vec) {
        int i = 2 * (int)threadIdx.x;
        for (; i + (kWgU - 1) * kWgStride + 1 < n; i += kWgU * kWgStride) {
            d2v xv[kWgU], yv[kWgU];
#pragma unroll
            for (int u = 0; u < kWgU; ++u) {
                xv[u] = *(const d2v *)(x + i + kWgStride * u);
                yv[u] = *(const d2v *)(y + i + kWgStride * u);
            }
#pragma unroll
            for (int u = 0; u < kWgU; ++u) {
                s = fma(xv[u].x, yv[u].x, s);
                s = fma(xv[u].y, yv[u].y, s);
            }
        }
        for (; i < n; i += kWgStride) {
            s = fma(x[i], y[i], s);
            if (i + 1 < n) s = fma(x[i + 1], y[i + 1], s);
        }
    } else {
        for_pairs(n, [&](int i) { s = fma(x[x0 + i * incx], y[y0 + i * incy], s); });
    }
    s = block_sum_wg(s, red);
    if (threadIdx.x == 0) {
        *result = s;
        publish(flag, seq);
    }
}

__device__ __forceinline__ void nrm2_wg(int n, const double *x, long incx, long x0, int vec,
                                        double *result, uint64_t *flag, uint64_t seq)
{
    __shared__ double red[kWgWaves];
    Acc3 a{0, 0, 0};
    if (vec) {
        int i = 2 * (int)threadIdx.x;
        for (; i + (kWgU - 1) * kWgStride + 1 < n; i += kWgU * kWgStride) {
            d2v xv[kWgU];
#pragma unroll
            for (int u = 0; u < kWgU; ++u) xv[u] = *(const d2v *)(x + i + kWgStride * u);
#pragma unroll
            for (int u = 0; u < kWgU; ++u) {
                acc3_add(a, xv[u].x);
                acc3_add(a, xv[u].y);
            }
        }
        for (; i < n; i += kWgStride) {
            acc3_add(a, x[i]);
            if (i + 1 < n) acc3_add(a, x[i + 1]);
        }
    } else {
        for_pairs(n, [&](int i) { acc3_add(a, x[x0 + i * incx]); });
    }
    a.sml = block_sum_wg(a.sml, red);
    a.med = block_sum_wg(a.med, red);
    a.big = block_sum_wg(a.big, red);
    if (threadIdx.x == 0) {
        *result = nrm2_finish(a.sml, a.med, a.big);
        publish(flag, seq);
    }
}

template <bool MIN>
__device__ __forceinline__ void iamax_wg(int n, const double *x, long incx, int vec, int *result,
                                         uint64_t *flag, uint64_t seq)
{
    __shared__ Cand red[kWgWaves];
    Cand c = cand_none<MIN>();
    auto take = [&](long i, double v) { c = better<MIN>(c, cand_of<MIN>(i, v)); };
    if (vec) {
        int i = 2 * (int)threadIdx.x;
        for (; i + (kWgU - 1) * kWgStride + 1 < n; i += kWgU * kWgStride) {
            d2v xv[kWgU];
#pragma unroll
            for (int u = 0; u < kWgU; ++u) xv[u] = *(const d2v *)(x + i + kWgStride * u);
#pragma unroll
            for (int u = 0; u < kWgU; ++u) {
                take(i + kWgStride * u, xv[u].x);
                take(i + kWgStride * u + 1, xv[u].y);
            }
        }
        for (; i < n; i += kWgStride) {
            take(i, x[i]);
            if (i + 1 < n) take(i + 1, x[i + 1]);
        }
    } else {
        for_pairs(n, [&](int i) { take(i, x[i * incx]); });
    }
    c = wave_best<MIN>(c);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        Cand t = red[0];
#pragma unroll
        for (int w = 1; w < kWgWaves; ++w) t = better<MIN>(t, red[w]);
        *result = (int)cand_result<MIN>(t, x);
        publish(flag, seq);
    }
}

__global__ __launch_bounds__(kWgThreads) void dot_small(int n, const double *x, long incx, long x0,
                                                  const double *y, long incy, long y0, int vec,
                                                  double *result, uint64_t *flag, uint64_t seq)
{
    dot_wg(n, x, incx, x0, y, incy, y0, vec, result, flag, seq);
}

__global__ __launch_bounds__(kWgThreads) void nrm2_small(int n, const double *x, long incx, long x0,
                                                   int vec, double *result, uint64_t *flag,
                                                   uint64_t seq)
{
    nrm2_wg(n, x, incx, x0, vec, result, flag, seq);
}

template <bool MIN>
__global__ __launch_bounds__(kWgThreads) void iamax_small(int n, const double *x, long incx, int vec,
                                                    int *result, uint64_t *flag, uint64_t seq)
{
    iamax_wg<MIN>(n, x, incx, vec, result, flag, seq);
}

inline long start_of(int n, int inc) { return inc < 0 ? -(long)(n - 1) * inc : 0; }
inline int aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

// ---- DSCAL / DCOPY / DSWAP / DROT: one strided element pair per thread -------
// (netlib orders: element i of a negative-increment vector sits at
// (n-1-i)*|inc| from its start, which x0/y0 and the signed increments encode)
enum PairOp { OP_SCAL, OP_COPY, OP_SWAP, OP_ROT };

template <int OP>
__global__ __launch_bounds__(256) void pair_kernel(int n, double *x, long incx, long x0, double *y,
                                                   long incy, long y0, double c, double sn)
{
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        double *px = x + x0 + i * incx;
        if (OP == OP_SCAL) {
            *px = c * *px;
            continue;
        }
        double *py = y + y0 + i * incy;
        if (OP == OP_COPY) {
            *py = *px;
        } else if (OP == OP_SWAP) {
            const double t = *px;
            *px = *py;
            *py = t;
        } else {
            const double xv = *px, yv = *py;
            *px = c * xv + sn * yv;
            *py = c * yv - sn * xv;
        }
    }
}

template <int OP>
int launch_pair(hipStream_t s, const char *name, int n, double *x, int incx, double *y, int incy,
                double c, double sn)
{
    int blocks = (int)(((long)n + 255) / 256);
    if (blocks > 16384) blocks = 16384;
    PROF_BEGIN(s, name);
    hipLaunchKernelGGL(pair_kernel<OP>, dim3(blocks), dim3(256), 0, s, n, x, (long)incx,
                       start_of(n, incx), y, (long)incy, start_of(n, incy), c, sn);
    PROF_END(s);
    return (int)hipGetLastError();
}

// ---- DROTM: (x_i, y_i) := H (x_i, y_i) with the modified-Givens H of the
// flag, in the netlib DROTM forms (FLAG < 0: full H; 0: unit diagonal;
// 1: h12 = 1, h21 = -1) ----------------------------------------------------------
template <int FLAG>
__global__ __launch_bounds__(256) void rotm_kernel(int n, double *x, long incx, long x0, double *y,
                                                   long incy, long y0, double h11, double h21,
                                                   double h12, double h22)
{
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        double *px = x + x0 + i * incx, *py = y + y0 + i * incy;
        const double w = *px, z = *py;
        if (FLAG < 0) {
            *px = w * h11 + z * h12;
            *py = w * h21 + z * h22;
        } else if (FLAG == 0) {
            *px = w + z * h12;
            *py = w * h21 + z;
        } else {
            *px = w * h11 + z;
            *py = -w + h22 * z;
        }
    }
}

// ---- DASUM: the DDOT reductions with |x_i| in place of x_i*y_i -----------------
__device__ __forceinline__ void asum_wg(int n, const double *x, long incx, int vec,
                                        double *result, uint64_t *flag, uint64_t seq)
{
    __shared__ double red[kWgWaves];
    double s = 0;
    if (vec) {
        int i = 2 * (int)threadIdx.x;
        for (; i + (kWgU - 1) * kWgStride + 1 < n; i += kWgU * kWgStride) {
            d2v xv[kWgU];
#pragma unroll
            for (int u = 0; u < kWgU; ++u) xv[u] = *(const d2v *)(x + i + kWgStride * u);
#pragma unroll
            for (int u = 0; u < kWgU; ++u) {
                s += fabs(xv[u].x);
                s += fabs(xv[u].y);
            }
        }
        for (; i < n; i += kWgStride) {
            s += fabs(x[i]);
            if (i + 1 < n) s += fabs(x[i + 1]);
        }
    } else {
        for_pairs(n, [&](int i) { s += fabs(x[i * incx]); });
    }
    s = block_sum_wg(s, red);
    if (threadIdx.x == 0) {
        *result = s;
        publish(flag, seq);
    }
}

__global__ __launch_bounds__(kWgThreads) void asum_small(int n, const double *x, long incx, int vec,
                                                   double *result, uint64_t *flag, uint64_t seq)
{
    asum_wg(n, x, incx, vec, result, flag, seq);
}

__global__ __launch_bounds__(256) void asum_kernel(int n, const double *x, long incx,
                                                   double *partials, unsigned *ticket,
                                                   double *result, uint64_t *flag, uint64_t seq)
{
    __shared__ double red[4];
    double s = 0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        s += fabs(x[i * incx]);
    s = block_sum(s, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
    if (!last_block(ticket)) return;
    double t = 0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += 256) t += partials[b];
    t = block_sum(t, red);
    if (threadIdx.x == 0) {
        *result = t;
        publish(flag, seq);
        __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---- resident reduction server -------------------------------------------------
// A kernel that stays on the device and serves the short (n <= kSmallN)
// reductions DDOT / DNRM2 / IDAMAX / IDAMIN / DASUM from a ring of request
// slots in pinned host memory, so a scalar-returning call costs no kernel
// launch: the caller fills a slot and publishes its sequence number; the
// server workgroup (512 threads), already polling that slot, sees it --
// the whole descriptor in the same read, each word tagged with the request's
// id, so one trip over the link both finds the request and brings it in --
// runs the same single-workgroup reduction as the launched kernel (identical
// bits) and publishes the result in the slot.  It serves a call only when the
// kernels that wrote its operands have ended (the caller's stream has no work
// queued, or the caller knows those writers completed): a launch queued
// behind such work starts as soon as it ends, which a hand-over cannot beat,
// and stream order needs no device-side wait.  Every wait of the server is bounded:
// after `idle` steady-clock ticks with nothing to do it exits, and a later
// request starts it again; a caller whose request was not served (the server
// exited first) takes the launched path.  So a server that shares a hardware
// queue with other work delays that work by at most the idle time, and
// nothing waits on it forever.  Descriptor words and flags are read with
// atomic loads (never through the scalar cache, which a persistent kernel
// would not see refreshed); results are written with vector stores and
// published with a system-scope release.
enum { RS_DOT = 0, RS_NRM2, RS_IAMAX, RS_IAMIN, RS_ASUM };
constexpr int kResSlots = 64;

// Descriptor words: bits 63..48 carry the request's tag, bits 47..0 the
// field.  The host stores each word whole (8-byte aligned: single-copy
// atomic on both sides), so a read that finds the expected tag in all five
// words has the request's whole descriptor, whatever the order in which the
// link delivered the words.  A slot is reused every kResSlots requests, and
// tags of requests less than 2^15 apart differ, so stale words never match.
// Device addresses fit in 48 bits (the host checks each one).
constexpr int kResWords = 5;
struct alignas(128) ResSlot {
    uint64_t w[8];           // [0] op | vec << 3 | n << 4, [1] x, [2] y, [3] incx, [4] incy
    double dres;             // results, written by the server (their own 64 B line)
    int64_t ires;
    uint64_t done;           // = request id + 1 once the result is valid
    uint64_t pad[5];
};
static_assert(sizeof(ResSlot) == 128, "one slot per 128 B");

__host__ __device__ inline uint64_t res_tag(unsigned long long id)
{
    return (uint64_t)(((id + 1) & 0x7fffu) | 0x8000u) << 48;  // never zero: a cleared slot never matches
}
constexpr uint64_t kResField = (uint64_t(1) << 48) - 1;

__global__ __launch_bounds__(kWgThreads, 4) void res_server(ResSlot *ring, unsigned long long first,
                                                   const int *stop, long long idle)
{
    __shared__ uint64_t d[kResWords];
    __shared__ int quit;
    for (unsigned long long id = first;; ++id) {
        ResSlot *sl = ring + id % kResSlots;
        if (threadIdx.x < 64) {  // wave 0 polls: lane i reads word i
            const int lane = threadIdx.x;
            const uint64_t tag = res_tag(id);
            int q = 0;
            uint64_t v = 0;
            const long long t0 = wall_clock64();
            for (unsigned spin = 0;; ++spin) {
                bool ok = true;
                if (lane < kResWords) {
                    v = __hip_atomic_load(sl->w + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    ok = (v & ~kResField) == tag;
                }
                if (__ballot(ok) == __ballot(1)) break;  // all five words carry the tag
                if ((spin & 15) == 0) {
                    int sv = 0;
                    if (lane == 0)
                        sv = __hip_atomic_load(stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) ||
                             wall_clock64() - t0 > idle;
                    if (__builtin_amdgcn_readfirstlane(sv)) {
                        q = 1;
                        break;
                    }
                }
                __builtin_amdgcn_s_sleep(2);
            }
            if (lane < kResWords) d[lane] = v & kResField;
            if (lane == 0) quit = q;
        }
        __syncthreads();
        if (quit) return;
        // see the operands' latest contents (their writers ended before the
        // caller's stream went idle)
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        const int op = (int)(d[0] & 7), vec = (int)((d[0] >> 3) & 1), n = (int)(d[0] >> 4);
        const double *x = (const double *)d[1], *y = (const double *)d[2];
        const long incx = (long)(int)(uint32_t)d[3], incy = (long)(int)(uint32_t)d[4];
        const long x0 = incx < 0 ? -(long)(n - 1) * incx : 0;
        const long y0 = incy < 0 ? -(long)(n - 1) * incy : 0;
        int *ires = (int *)&sl->ires;
        switch (op) {
        case RS_DOT: dot_wg(n, x, incx, x0, y, incy, y0, vec, &sl->dres, &sl->done, id + 1); break;
        case RS_NRM2: nrm2_wg(n, x, incx, x0, vec, &sl->dres, &sl->done, id + 1); break;
        case RS_IAMAX: iamax_wg<false>(n, x, incx, vec, ires, &sl->done, id + 1); break;
        case RS_IAMIN: iamax_wg<true>(n, x, incx, vec, ires, &sl->done, id + 1); break;
        default: asum_wg(n, x, incx, vec, &sl->dres, &sl->done, id + 1); break;
        }
        __syncthreads();
    }
}

// Host side.  One request in flight at a time (a caller that finds another
// request in flight takes the launched path); ids never repeat, so a slot's
// `done` names exactly one request.
struct ResServer {
    std::mutex mu;
    bool enabled = false;   // gpublas_dev_resident_start() called, not stopped
    bool running = false;   // a server kernel is on the device
    ResSlot *ring = nullptr;  // pinned host memory
    int *stop = nullptr;      // pinned host memory
    hipStream_t stream = nullptr;
    uint64_t next = 0;        // next request id
    long long idle = 0;       // steady-clock ticks of idle before the server exits
    uint64_t idle_ns = 0;     // the same in host nanoseconds
    uint64_t last_ns = 0;     // host time the last request ended
    bool closed = false;      // process exit has begun: no new start
    bool forked = false;      // this is a fork child: the server is the parent's
};
ResServer g_res;

// At exit: stop the server and release its queue while the runtime is whole
// (not in a fork child, whose runtime is not the parent's).
void res_stop_at_exit()
{
    if (g_res.forked) return;
    gpublas_dev_resident_stop();
    std::lock_guard<std::mutex> g(g_res.mu);
    g_res.closed = true;
    if (g_res.stream) (void)hipStreamDestroy(g_res.stream);
    g_res.stream = nullptr;
}

uint64_t res_now_ns()
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (uint64_t)t.tv_sec * 1000000000ull + (uint64_t)t.tv_nsec;
}

// (mu held) make sure a server is running.  While requests come faster than
// half its idle time it is known to be alive; after a longer gap the stream
// is queried, and an exited server is launched again.  (One that exits in
// the moment between the query and the request is noticed by that request's
// wait, which then takes the launched path.)
int res_ensure_locked()
{
    if (g_res.running && res_now_ns() - g_res.last_ns < g_res.idle_ns / 2) return 0;
    if (g_res.running) {
        const hipError_t q = hipStreamQuery(g_res.stream);
        if (q == hipErrorNotReady) return 0;
        if (q != hipSuccess) return (int)q;
        g_res.running = false;
    }
    const hipError_t q = hipStreamSynchronize(g_res.stream);
    if (q != hipSuccess) return (int)q;
    *(volatile int *)g_res.stop = 0;
    hipLaunchKernelGGL(res_server, dim3(1), dim3(kWgThreads), 0, g_res.stream, g_res.ring,
                       (unsigned long long)g_res.next, (const int *)g_res.stop, g_res.idle);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    g_res.running = true;
    return 0;
}

}  // namespace

void res_forget_after_fork()
{
    new (&g_res.mu) std::mutex();  // possibly held by a parent thread at the fork
    g_res.forked = true;
    g_res.closed = true;
    g_res.enabled = false;
    g_res.running = false;
}

extern "C" {

int gpublas_dev_resident_start(void)
{
    if (gpublas_dev_init() != 0) return (int)hipErrorNoDevice;
    std::lock_guard<std::mutex> g(g_res.mu);
    if (g_res.enabled) return 0;
    if (g_res.closed) return (int)hipErrorNotReady;
    if (!g_res.ring) {
        // a queue of its own where the runtime provides one (CU-masked streams)
        uint32_t mask[8];
        for (uint32_t &m : mask) m = 0xffffffffu;
        if (hipExtStreamCreateWithCUMask(&g_res.stream, 8, mask) != hipSuccess &&
            hipStreamCreateWithFlags(&g_res.stream, hipStreamNonBlocking) != hipSuccess)
            return (int)hipErrorOutOfMemory;
        if (hipHostMalloc((void **)&g_res.ring, kResSlots * sizeof(ResSlot), hipHostMallocCoherent) !=
                hipSuccess ||
            hipHostMalloc((void **)&g_res.stop, 64, hipHostMallocCoherent) != hipSuccess)
            return (int)hipErrorOutOfMemory;
        memset((void *)g_res.ring, 0, kResSlots * sizeof(ResSlot));
        int rate_khz = 0;
        if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, 0) != hipSuccess ||
            rate_khz <= 0)
            rate_khz = 100000;
        g_res.idle = (long long)rate_khz * 2;  // 2 ms
        g_res.idle_ns = 2000000;
        atexit(res_stop_at_exit);
    }
    g_res.enabled = true;
    return 0;
}

void gpublas_dev_resident_stop(void)
{
    std::lock_guard<std::mutex> g(g_res.mu);
    g_res.enabled = false;
    if (!g_res.running) return;
    __atomic_store_n(g_res.stop, 1, __ATOMIC_RELEASE);
    (void)hipStreamSynchronize(g_res.stream);
    g_res.running = false;
}

int gpublas_dev_resident_active(void) { return __atomic_load_n(&g_res.enabled, __ATOMIC_ACQUIRE); }

int gpublas_dev_resident_reduce(void *stream, int op, int n, const double *d_x, int incx,
                                const double *d_y, int incy, int operands_ready, double *dres,
                                int *ires)
{
    if (!__atomic_load_n(&g_res.enabled, __ATOMIC_ACQUIRE)) return -1;
    if (n < 1 || n > kSmallN || op < RS_DOT || op > RS_ASUM) return -1;
    if ((op == RS_IAMAX || op == RS_IAMIN || op == RS_ASUM) && incx <= 0) return -1;
    if (((uintptr_t)d_x | (uintptr_t)d_y) & ~(uintptr_t)kResField) return -1;  // beyond 48 bits
    hipStream_t s = pick_stream(stream);
    DevCtx *c = ctx_of(s);
    if (!c) return -1;
    std::unique_lock<std::mutex> lk(g_res.mu, std::try_to_lock);
    if (!lk.owns_lock() || !g_res.enabled) return -1;  // another request in flight
    if (res_ensure_locked() != 0) return -1;
    // Unless the caller knows the operands' writers have ended, only when the
    // caller's stream has drained: behind queued work a launch starts as soon
    // as that work ends, which no hand-over to the server beats.
    if (!operands_ready) {
        const hipError_t q = hipStreamQuery(s);
        if (q == hipErrorNotReady) return -1;
        if (q != hipSuccess) return (int)q;
    }
    const uint64_t id = g_res.next++;
    ResSlot *sl = g_res.ring + id % kResSlots;
    const int vy = op == RS_DOT ? (((uintptr_t)d_y & 15) == 0 && incy == 1) : 1;
    const int vec = incx == 1 && ((uintptr_t)d_x & 15) == 0 && vy;
    const uint64_t tag = res_tag(id);
    __atomic_store_n(&sl->w[1], tag | (uint64_t)(uintptr_t)d_x, __ATOMIC_RELAXED);
    __atomic_store_n(&sl->w[2], tag | (uint64_t)(uintptr_t)d_y, __ATOMIC_RELAXED);
    __atomic_store_n(&sl->w[3], tag | (uint32_t)incx, __ATOMIC_RELAXED);
    __atomic_store_n(&sl->w[4], tag | (uint32_t)incy, __ATOMIC_RELAXED);
    __atomic_store_n(&sl->w[0], tag | (uint64_t)op | (uint64_t)vec << 3 | (uint64_t)n << 4,
                     __ATOMIC_RELEASE);
    for (unsigned spin = 1;; ++spin) {
        if (__atomic_load_n(&sl->done, __ATOMIC_ACQUIRE) == id + 1) break;
        if ((spin & 1023) == 0) {
            const hipError_t e = hipStreamQuery(g_res.stream);
            if (e == hipErrorNotReady) continue;
            // the server exited (idle time, or an error): served after all?
            g_res.running = false;
            if (__atomic_load_n(&sl->done, __ATOMIC_ACQUIRE) == id + 1) break;
            return -1;  // not served: the caller takes the launched path
        }
    }
    if (dres) *dres = *(volatile double *)&sl->dres;
    if (ires) *ires = (int)*(volatile int64_t *)&sl->ires;
    g_res.last_ns = res_now_ns();
    return 0;
}

int gpublas_dev_daxpy(void *stream, int n, double alpha, const double *d_x, int incx,
                      double *d_y, int incy)
{
    if (n <= 0 || alpha == 0.0) return 0;
    hipStream_t s = pick_stream(stream);
    const bool unit = incx == 1 && incy == 1 && (n % 2 == 0) && (((uintptr_t)d_x & 15) == 0) &&
                      (((uintptr_t)d_y & 15) == 0);
    PROF_BEGIN(s, "daxpy");
    if (unit) {
        long n2 = n / 2;
        int blocks = (int)((n2 + 255) / 256);
        if (blocks > 16384) blocks = 16384;
        hipLaunchKernelGGL(axpy_unit_v2, dim3(blocks), dim3(256), 0, s, n2, alpha,
                           (const double2 *)d_x, (double2 *)d_y);
    } else {
        int blocks = (int)(((long)n + 255) / 256);
        if (blocks > 16384) blocks = 16384;
        hipLaunchKernelGGL(axpy_strided, dim3(blocks), dim3(256), 0, s, n, alpha, d_x, (long)incx,
                           start_of(n, incx), d_y, (long)incy, start_of(n, incy));
    }
    PROF_END(s);
    return (int)hipGetLastError();
}

int gpublas_dev_ddot(void *stream, int n, const double *d_x, int incx, const double *d_y, int incy,
                     double *h_result)
{
    hipStream_t s = pick_stream(stream);
    if (n <= 0) {
        *h_result = 0.0;
        return 0;
    }
    const Pub pb = pub_for(h_result);
    if (n <= kSmallN) {
        const int vec = incx == 1 && incy == 1 && aligned16(d_x) && aligned16(d_y);
        PROF_BEGIN(s, "ddot");
        hipLaunchKernelGGL(dot_small, dim3(1), dim3(kWgThreads), 0, s, n, d_x, (long)incx,
                           start_of(n, incx), d_y, (long)incy, start_of(n, incy), vec, h_result,
                           pb.flag, pb.seq);
        PROF_END(s);
        return (int)hipGetLastError();
    }
    RedSlot r = next_slot(s);
    if (!r.partials) return (int)hipErrorOutOfMemory;
    int blocks = red_blocks(n);
    PROF_BEGIN(s, "ddot");
    hipLaunchKernelGGL(dot_kernel, dim3(blocks), dim3(256), 0, s, n, d_x, (long)incx,
                       start_of(n, incx), d_y, (long)incy, start_of(n, incy),
                       (double *)r.partials, r.ticket, h_result, pb.flag, pb.seq);
    PROF_END(s);
    return (int)hipGetLastError();
}

int gpublas_dev_dnrm2(void *stream, int n, const double *d_x, int incx, double *h_result)
{
    hipStream_t s = pick_stream(stream);
    if (n <= 0) {
        *h_result = 0.0;
        return 0;
    }
    const Pub pb = pub_for(h_result);
    if (n <= kSmallN) {
        const int vec = incx == 1 && aligned16(d_x);
        PROF_BEGIN(s, "dnrm2");
        hipLaunchKernelGGL(nrm2_small, dim3(1), dim3(kWgThreads), 0, s, n, d_x, (long)incx,
                           start_of(n, incx), vec, h_result, pb.flag, pb.seq);
        PROF_END(s);
        return (int)hipGetLastError();
    }
    RedSlot r = next_slot(s);
    if (!r.partials) return (int)hipErrorOutOfMemory;
    int blocks = red_blocks(n);
    if (blocks > kMaxBlocks / 2) blocks = kMaxBlocks / 2;  // 3 partials per block
    PROF_BEGIN(s, "dnrm2");
    hipLaunchKernelGGL(nrm2_kernel, dim3(blocks), dim3(256), 0, s, n, d_x, (long)incx,
                       start_of(n, incx), (double *)r.partials, r.ticket, h_result, pb.flag,
                       pb.seq);
    PROF_END(s);
    return (int)hipGetLastError();
}

}  // extern "C"

namespace {

template <bool MIN>
int launch_iamax(void *stream, int n, const double *d_x, int incx, int *h_result)
{
    hipStream_t s = pick_stream(stream);
    if (n < 1 || incx <= 0) {
        *h_result = 0;
        return 0;
    }
    const char *name = MIN ? "idamin" : "idamax";
    const Pub pb = pub_for(h_result);
    if (n <= kSmallN) {
        const int vec = incx == 1 && aligned16(d_x);
        PROF_BEGIN(s, name);
        hipLaunchKernelGGL(iamax_small<MIN>, dim3(1), dim3(kWgThreads), 0, s, n, d_x, (long)incx, vec,
                           h_result, pb.flag, pb.seq);
        PROF_END(s);
        return (int)hipGetLastError();
    }
    RedSlot r = next_slot(s);
    if (!r.partials) return (int)hipErrorOutOfMemory;
    int blocks = red_blocks(n);
    PROF_BEGIN(s, name);
    hipLaunchKernelGGL(iamax_kernel<MIN>, dim3(blocks), dim3(256), 0, s, n, d_x, (long)incx,
                       (double *)r.partials, r.ticket, h_result, pb.flag, pb.seq);
    PROF_END(s);
    return (int)hipGetLastError();
}

}  // namespace

extern "C" {

int gpublas_dev_idamax(void *stream, int n, const double *d_x, int incx, int *h_result)
{
    return launch_iamax<false>(stream, n, d_x, incx, h_result);
}

int gpublas_dev_idamin(void *stream, int n, const double *d_x, int incx, int *h_result)
{
    return launch_iamax<true>(stream, n, d_x, incx, h_result);
}

int gpublas_dev_drotm(void *stream, int n, double *d_x, int incx, double *d_y, int incy,
                      const double *h_param)
{
    const double flag = h_param[0];
    if (n <= 0 || flag == -2.0) return 0;
    hipStream_t s = pick_stream(stream);
    int blocks = (int)(((long)n + 255) / 256);
    if (blocks > 16384) blocks = 16384;
    const long x0 = start_of(n, incx), y0 = start_of(n, incy);
    PROF_BEGIN(s, "drotm");
#define GB_RM(F)                                                                                   \
    hipLaunchKernelGGL(rotm_kernel<F>, dim3(blocks), dim3(256), 0, s, n, d_x, (long)incx, x0, d_y, \
                       (long)incy, y0, h_param[1], h_param[2], h_param[3], h_param[4])
    if (flag < 0) GB_RM(-1);
    else if (flag == 0) GB_RM(0);
    else GB_RM(1);
#undef GB_RM
    PROF_END(s);
    return (int)hipGetLastError();
}

int gpublas_dev_dscal(void *stream, int n, double alpha, double *d_x, int incx)
{
    if (n <= 0 || incx <= 0) return 0;  // netlib DSCAL: no-op for incx <= 0
    return launch_pair<OP_SCAL>(pick_stream(stream), "dscal", n, d_x, incx, nullptr, 1, alpha, 0.0);
}

int gpublas_dev_dcopy(void *stream, int n, const double *d_x, int incx, double *d_y, int incy)
{
    if (n <= 0) return 0;
    return launch_pair<OP_COPY>(pick_stream(stream), "dcopy", n, (double *)d_x, incx, d_y, incy,
                                0.0, 0.0);
}

int gpublas_dev_dswap(void *stream, int n, double *d_x, int incx, double *d_y, int incy)
{
    if (n <= 0) return 0;
    return launch_pair<OP_SWAP>(pick_stream(stream), "dswap", n, d_x, incx, d_y, incy, 0.0, 0.0);
}

int gpublas_dev_drot(void *stream, int n, double *d_x, int incx, double *d_y, int incy, double c,
                     double s)
{
    if (n <= 0) return 0;
    return launch_pair<OP_ROT>(pick_stream(stream), "drot", n, d_x, incx, d_y, incy, c, s);
}

int gpublas_dev_dasum(void *stream, int n, const double *d_x, int incx, double *h_result)
{
    hipStream_t s = pick_stream(stream);
    if (n <= 0 || incx <= 0) {  // netlib DASUM returns 0
        *h_result = 0.0;
        return 0;
    }
    const Pub pb = pub_for(h_result);
    if (n <= kSmallN) {
        const int vec = incx == 1 && aligned16(d_x);
        PROF_BEGIN(s, "dasum");
        hipLaunchKernelGGL(asum_small, dim3(1), dim3(kWgThreads), 0, s, n, d_x, (long)incx, vec,
                           h_result, pb.flag, pb.seq);
        PROF_END(s);
        return (int)hipGetLastError();
    }
    RedSlot r = next_slot(s);
    if (!r.partials) return (int)hipErrorOutOfMemory;
    int blocks = red_blocks(n);
    PROF_BEGIN(s, "dasum");
    hipLaunchKernelGGL(asum_kernel, dim3(blocks), dim3(256), 0, s, n, d_x, (long)incx,
                       (double *)r.partials, r.ticket, h_result, pb.flag, pb.seq);
    PROF_END(s);
    return (int)hipGetLastError();
}

}  // extern "C"
