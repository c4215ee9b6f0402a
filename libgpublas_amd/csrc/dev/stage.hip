// stage.hip -- host-link copies of operands in pageable host memory.
//
// A BLAS call whose operand is not resident copies it to HBM and its result
// back (the reference does this per call with cudaMemcpy, blas_level3/
// gemm.cc:46-83).  HIP's pageable copies reach ~30-36 GB/s host-to-device and
// only ~17 GB/s device-to-host on this machine, and a pageable 1 MiB copy can
// take hundreds of microseconds; pinned copies run at ~57 GB/s both ways and
// 4 host threads move 90-110 GB/s between pageable and pinned memory.  So
// the transfer is split into pieces of at most 4 MiB that go through a ring
// of pinned slots: host threads pack a piece into a slot while the DMA engine
// moves the previous one.
//
//   h2d: (from 1 MiB; smaller uploads are direct pageable copies)
//        pack piece i into slot s (after the slot's last DMA finished, an
//        event), enqueue the slot's DMA on the call's stream, record the
//        slot's event.  Returns with the DMAs in flight: the kernel that
//        follows on the stream waits for them; the user's buffer is no
//        longer read.
//   d2h: enqueue DMAs for the first pieces, then for each piece wait for its
//        event, unpack it into the user's buffer and reuse the slot for the
//        next piece.  Returns with the data in host memory.
//
// Column-major blocks with leading dimensions keep that layout: a piece is a
// group of whole columns (packed densely in the slot, hipMemcpy2DAsync between
// slot and device), or a byte range of one column when a column exceeds a
// slot.  Only the block's own elements are written on either side.
//
// One staged transfer at a time (the host link is shared); concurrent callers
// queue on a mutex.  The memcpy pool is 3 worker threads plus the caller; they
// spin briefly between pieces and sleep on a condition variable otherwise.
#include "devcommon.hpp"
#include <pthread.h>
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>

namespace {

constexpr size_t kSlotBytes = 4u << 20;
constexpr int kSlots = 4;
constexpr size_t kPartMin = 512u << 10;  // smallest memcpy share of one thread
constexpr int kWorkers = 3;
// Uploads below this go straight from pageable memory: measured per call
// (3 uploads + a kernel + a download), HIP's own small pageable copies are
// 15-40 us faster than slot copies up to 512 KiB per operand; from 1 MiB its
// pageable path is 8x slower (363 us vs 47 us).  Downloads are staged at
// every size (faster from 512 B up).
constexpr size_t kDirectUpload = 1u << 20;

// A 2-D byte region: height rows of width bytes, with pitches.
struct Region {
    char *dst;
    const char *src;
    size_t dpitch, spitch, width, height;
};

inline void cpu_relax() { __builtin_ia32_pause(); }

// Copy part p of P of a region (rows split when there are enough, else bytes).
void copy_part(const Region &r, unsigned p, unsigned P)
{
    if (r.height >= P) {
        const size_t h0 = r.height * p / P, h1 = r.height * (p + 1) / P;
        for (size_t h = h0; h < h1; ++h)
            memcpy(r.dst + h * r.dpitch, r.src + h * r.spitch, r.width);
    } else {
        // P > height: P is a multiple of height (see parts_for); split each row
        const unsigned per = P / (unsigned)r.height;
        const size_t h = p / per, q = p % per;
        const size_t b0 = r.width * q / per, b1 = r.width * (q + 1) / per;
        memcpy(r.dst + h * r.dpitch + b0, r.src + h * r.spitch + b0, b1 - b0);
    }
}

unsigned parts_for(const Region &r)
{
    const size_t bytes = r.width * r.height;
    const size_t P = std::min<size_t>(kWorkers + 1, std::max<size_t>(1, bytes / kPartMin));
    if (P <= 1 || r.height >= P) return (unsigned)P;
    // fewer rows than parts: every row split into the same number of shares
    return (unsigned)((P / r.height) * r.height);
}

// Ticket: generation (48 bits) | parts (8 bits) | next part (8 bits).
struct CopyPool {
    std::atomic<uint64_t> ticket{0};
    std::atomic<unsigned> done{0};
    Region job{};
    std::mutex mu;
    std::condition_variable cv;
    bool started = false;

    static unsigned parts_of(uint64_t t) { return (unsigned)((t >> 8) & 0xff); }
    static unsigned next_of(uint64_t t) { return (unsigned)(t & 0xff); }

    // claim and run one part of the current job; false if none is left
    bool work_one()
    {
        uint64_t t = ticket.load(std::memory_order_acquire);
        for (;;) {
            const unsigned P = parts_of(t), i = next_of(t);
            if (i >= P) return false;
            if (ticket.compare_exchange_weak(t, t + 1, std::memory_order_acq_rel,
                                             std::memory_order_acquire)) {
                copy_part(job, i, P);
                done.fetch_add(1, std::memory_order_release);
                return true;
            }
        }
    }

    void worker()
    {
        for (;;) {
            if (work_one()) continue;
            bool got = false;
            for (int spin = 0; spin < 20000 && !got; ++spin) {
                const uint64_t t = ticket.load(std::memory_order_acquire);
                got = next_of(t) < parts_of(t);
                if (!got) cpu_relax();
            }
            if (got) continue;
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] {
                const uint64_t t = ticket.load(std::memory_order_acquire);
                return next_of(t) < parts_of(t);
            });
        }
    }

    void start()
    {
        if (started) return;
        started = true;
        for (int w = 0; w < kWorkers; ++w) std::thread([this] { worker(); }).detach();
    }

    // run a region copy on the pool and the calling thread; returns when done
    void run(const Region &r)
    {
        if (!r.width || !r.height) return;
        const unsigned P = parts_for(r);
        if (P <= 1) {
            copy_part(r, 0, 1);
            return;
        }
        start();
        job = r;
        done.store(0, std::memory_order_relaxed);
        const uint64_t gen = (ticket.load(std::memory_order_relaxed) >> 16) + 1;
        ticket.store((gen << 16) | ((uint64_t)P << 8), std::memory_order_release);
        { std::lock_guard<std::mutex> lk(mu); }
        cv.notify_all();
        while (work_one()) {
        }
        while (done.load(std::memory_order_acquire) < P) cpu_relax();
    }
};

struct Stager {
    std::mutex mu;
    bool tried = false, ok = false;
    char *slot[kSlots] = {};
    hipEvent_t ev[kSlots] = {};
    bool live[kSlots] = {};
    int next = 0;
    CopyPool pool;

    bool init()
    {
        if (tried) return ok;
        tried = true;
        char *base = nullptr;
        if (hipHostMalloc((void **)&base, kSlotBytes * kSlots, hipHostMallocDefault) != hipSuccess)
            return false;
        for (int i = 0; i < kSlots; ++i) {
            slot[i] = base + i * kSlotBytes;
            if (hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess) return false;
        }
        ok = true;
        return true;
    }

    // wait until slot s is free (its last DMA finished).  Polled: a sleeping
    // wait wakes late when other threads (a host BLAS's) keep the cores busy.
    hipError_t drain(int s)
    {
        if (!live[s]) return hipSuccess;
        live[s] = false;
        for (;;) {
            const hipError_t q = hipEventQuery(ev[s]);
            if (q != hipErrorNotReady) return q;
            cpu_relax();
        }
    }
};

// Lives until the process ends and is never destroyed: the pool's workers
// sleep on its condition variable, and destroying a condition variable with
// waiters blocks forever in glibc (it waits for them to wake).
Stager &g_stage = *new Stager();

// After fork() the child has none of the pool's threads: start afresh.
void stage_atfork_child()
{
    new (&g_stage.pool.mu) std::mutex();
    new (&g_stage.pool.cv) std::condition_variable();
    g_stage.pool.started = false;
    g_stage.pool.ticket.store(0);
    new (&g_stage.mu) std::mutex();
}
struct AtFork {
    AtFork() { pthread_atfork(nullptr, nullptr, stage_atfork_child); }
} g_atfork;

// A piece: cols columns of width bytes (host pitch hp, device pitch dp), or a
// byte range of one column (cols == 1).
struct Piece {
    char *h;
    char *d;
    size_t hp, dp, width, cols;
};

// Split an m x n block (element size es) into pieces of at most one slot.
template <typename F>
hipError_t for_pieces(char *h, size_t ldh, char *d, size_t ldd, size_t m, size_t n, size_t es, F f)
{
    const size_t colb = m * es;
    if (ldh == m && ldd == m) {
        // contiguous on both sides: byte ranges
        const size_t total = colb * n;
        for (size_t off = 0; off < total; off += kSlotBytes) {
            const size_t len = std::min(kSlotBytes, total - off);
            hipError_t e = f(Piece{h + off, d + off, len, len, len, 1});
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    if (colb <= kSlotBytes) {
        const size_t per = kSlotBytes / colb;
        for (size_t j = 0; j < n; j += per) {
            const size_t c = std::min(per, n - j);
            hipError_t e = f(Piece{h + j * ldh * es, d + j * ldd * es, ldh * es, ldd * es, colb, c});
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    for (size_t j = 0; j < n; ++j)
        for (size_t off = 0; off < colb; off += kSlotBytes) {
            const size_t len = std::min(kSlotBytes, colb - off);
            hipError_t e = f(Piece{h + j * ldh * es + off, d + j * ldd * es + off, len, len, len, 1});
            if (e != hipSuccess) return e;
        }
    return hipSuccess;
}

hipError_t dma_h2d(const Piece &p, const char *slot, hipStream_t s)
{
    if (p.cols == 1) return hipMemcpyAsync(p.d, slot, p.width, hipMemcpyHostToDevice, s);
    return hipMemcpy2DAsync(p.d, p.dp, slot, p.width, p.width, p.cols, hipMemcpyHostToDevice, s);
}

hipError_t dma_d2h(const Piece &p, char *slot, hipStream_t s)
{
    if (p.cols == 1) return hipMemcpyAsync(slot, p.d, p.width, hipMemcpyDeviceToHost, s);
    return hipMemcpy2DAsync(slot, p.width, p.d, p.dp, p.width, p.cols, hipMemcpyDeviceToHost, s);
}

}  // namespace

extern "C" int gpublas_dev_h2d_staged(void *d_dst, size_t ldd, const void *h_src, size_t ldh,
                                      size_t m, size_t n, size_t es, void *stream)
{
    if (!m || !n) return 0;
    hipStream_t s = pick_stream(stream);
    if (m * n * es < kDirectUpload) {
        if (ldh == m && ldd == m)
            return (int)hipMemcpyAsync(d_dst, h_src, m * n * es, hipMemcpyHostToDevice, s);
        return (int)hipMemcpy2DAsync(d_dst, ldd * es, h_src, ldh * es, m * es, n,
                                     hipMemcpyHostToDevice, s);
    }
    std::lock_guard<std::mutex> lk(g_stage.mu);
    if (!g_stage.init()) {
        if (ldh == m && ldd == m)
            return (int)hipMemcpyAsync(d_dst, h_src, m * n * es, hipMemcpyHostToDevice, s);
        return (int)hipMemcpy2DAsync(d_dst, ldd * es, h_src, ldh * es, m * es, n,
                                     hipMemcpyHostToDevice, s);
    }
    return (int)for_pieces((char *)h_src, ldh, (char *)d_dst, ldd, m, n, es,
                           [&](const Piece &p) -> hipError_t {
                               const int k = g_stage.next;
                               g_stage.next = (k + 1) % kSlots;
                               hipError_t e = g_stage.drain(k);
                               if (e != hipSuccess) return e;
                               g_stage.pool.run(Region{g_stage.slot[k], p.h, p.width, p.hp,
                                                       p.width, p.cols});
                               e = dma_h2d(p, g_stage.slot[k], s);
                               if (e != hipSuccess) return e;
                               e = hipEventRecord(g_stage.ev[k], s);
                               g_stage.live[k] = e == hipSuccess;
                               return e;
                           });
}

extern "C" int gpublas_dev_d2h_staged(void *h_dst, size_t ldh, const void *d_src, size_t ldd,
                                      size_t m, size_t n, size_t es, void *stream)
{
    if (!m || !n) return 0;
    hipStream_t s = pick_stream(stream);
    std::lock_guard<std::mutex> lk(g_stage.mu);
    if (!g_stage.init()) {
        if (ldh == m && ldd == m)
            return (int)hipMemcpyAsync(h_dst, d_src, m * n * es, hipMemcpyDeviceToHost, s);
        return (int)hipMemcpy2DAsync(h_dst, ldh * es, d_src, ldd * es, m * es, n,
                                     hipMemcpyDeviceToHost, s);
    }
    // pieces in flight: slot k holds piece q[k]; unpack in order, refill
    Piece q[kSlots];
    int head = 0, count = 0;  // ring of in-flight pieces, oldest at head
    auto finish_oldest = [&]() -> hipError_t {
        const int k = (g_stage.next - count + 2 * kSlots) % kSlots;
        hipError_t e = g_stage.drain(k);
        if (e != hipSuccess) return e;
        const Piece &p = q[head];
        g_stage.pool.run(Region{p.h, g_stage.slot[k], p.hp, p.width, p.width, p.cols});
        head = (head + 1) % kSlots;
        count--;
        return hipSuccess;
    };
    hipError_t e = for_pieces((char *)h_dst, ldh, (char *)d_src, ldd, m, n, es,
                              [&](const Piece &p) -> hipError_t {
                                  if (count == kSlots) {
                                      hipError_t e2 = finish_oldest();
                                      if (e2 != hipSuccess) return e2;
                                  }
                                  const int k = g_stage.next;
                                  hipError_t e2 = g_stage.drain(k);  // an earlier h2d's DMA
                                  if (e2 != hipSuccess) return e2;
                                  e2 = dma_d2h(p, g_stage.slot[k], s);
                                  if (e2 != hipSuccess) return e2;
                                  e2 = hipEventRecord(g_stage.ev[k], s);
                                  if (e2 != hipSuccess) return e2;
                                  g_stage.live[k] = true;
                                  g_stage.next = (k + 1) % kSlots;
                                  q[(head + count) % kSlots] = p;
                                  count++;
                                  return hipSuccess;
                              });
    while (e == hipSuccess && count > 0) e = finish_oldest();
    return (int)e;
}
