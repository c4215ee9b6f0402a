// devcommon.hpp -- shared pieces of the gfx950 device layer.
//
// The device layer is a thin C ABI (include/gpublas_dev.h) over hand-written
// HIP kernels.  Every entry point takes device addresses, a stream and plain
// sizes; none of them allocates on the host's behalf except through the
// caching pool in runtime.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#define GPUBLAS_WAVE 64

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// Per-thread device context: the stream every call of this host thread is
// ordered on, the reduction workspace that lives on that stream, and a
// pinned slot that kernels write scalar results into.
struct DevCtx {
    hipStream_t stream;
    void *ws;              // device workspace (reduction partials)
    size_t ws_bytes;
    unsigned *tickets;     // device ticket counters, kept zero between launches
    void *pinned;          // 256 B pinned host slot: [0] scalar result, [8] publish flag
    uint64_t submitted;    // calls marked on this stream (gpublas_dev_stream_mark)
    uint64_t completed;    // calls known complete (a sync observed them)
    uint64_t scalar_seq;   // last sequence number published into the pinned slot
};
// Context owning `stream` (contexts and their streams live for the process).
DevCtx *ctx_of(hipStream_t stream);

// Returns the calling thread's context (takes a pooled one, or creates it, on
// first use; returns it to the pool at thread exit -- streams are never
// destroyed, so a stream handle recorded by the host layer stays valid).  The
// stream argument of the C ABI may be NULL (use this thread's stream) or a
// stream of the caller's.
DevCtx *dev_ctx();
void dev_release_after(void *d_ptr, hipStream_t s);  // pool release ordered behind stream s

// Scratch memory (partial sums, tickets) of one call whose kernels run on
// stream s.  On the calling thread's own stream it is the thread's workspace:
// only that stream ever uses it, so growing it after a sync of that stream is
// safe.  On any other stream it is a pool block, released behind s (once the
// work queued on s so far completes) when the Scratch goes out of scope --
// after the call has enqueued its launches.  Two calls on two streams thus
// never share partials, and no queued kernel loses its workspace.
struct Scratch {
    void *p = nullptr;
    hipStream_t s = nullptr;
    bool pooled = false;
    Scratch() = default;
    Scratch(const Scratch &) = delete;
    Scratch &operator=(const Scratch &) = delete;
    Scratch(Scratch &&o) noexcept : p(o.p), s(o.s), pooled(o.pooled) { o.pooled = false; }
    Scratch &operator=(Scratch &&o) noexcept
    {
        if (this != &o) {
            if (pooled && p) dev_release_after(p, s);
            p = o.p, s = o.s, pooled = o.pooled;
            o.pooled = false;
        }
        return *this;
    }
    ~Scratch()
    {
        if (pooled && p) dev_release_after(p, s);
    }
};
Scratch dev_scratch(size_t bytes, hipStream_t s);

static constexpr int kTicketSlots = 4096;  // ticket counters per context

// Records, per device, the address of the fixed scratch array of the
// allocation-free DGEMV (blas2.hip); the library's init calls it once with
// the device current.  Returns 0 or a HIP error code.
int noalloc_scratch_setup(int device);
bool dev_ready();  // gpublas_dev_init() has run and found a gfx950 device (no init here)
// (blas1.hip) drop the resident server's state in a fork child, no runtime call
void res_forget_after_fork();

// Kernel timing (HIP events on the launch stream), enabled by
// gpublas_dev_prof_enable.  Wrap each launch as
//   PROF_BEGIN(stream, "name"); hipLaunchKernelGGL(...); PROF_END(stream);
void prof_begin(hipStream_t s, const char *name);
void prof_end(hipStream_t s);
extern int g_prof_on;
#define PROF_BEGIN(s, name) do { if (g_prof_on) prof_begin((s), (name)); } while (0)
#define PROF_END(s) do { if (g_prof_on) prof_end((s)); } while (0)

static inline hipStream_t pick_stream(void *s) { return s ? (hipStream_t)s : dev_ctx()->stream; }

static inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

// Wave-level reductions with DPP/ds_swizzle via __shfl_xor (wave64).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// s_waitcnt immediate: wait until at most n vector-memory operations of this
// wave are outstanding (LDS and export counters left at their maxima)
__host__ __device__ constexpr int vm_wait(int n)
{
    return (n & 15) | ((n >> 4) << 14) | (7 << 4) | (15 << 8);
}

// Workgroup barrier that waits only for this wave's LDS operations
// (lgkmcnt 0): global loads stay in flight across it, where the fence of
// __syncthreads would drain them.  For barriers that order LDS traffic alone.
// The scheduling barriers keep all code on its side: arithmetic hoisted above
// it onto registers whose loads were issued just before would make the
// compiler wait for those loads here.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(vm_wait(63) & ~(15 << 8));
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_sched_barrier(0);
}
