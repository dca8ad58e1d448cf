// hpdex_stage.cpp -- copies between caller (pageable) memory and the device
// through pinned staging buffers.
//
// A copy from pageable memory is bounded by the runtime's own staging (27.9
// GB/s measured on one MI355X host against 57.6 GB/s from pinned memory).
// Here a copy is cut into chunks: host threads fill a pinned buffer of a ring
// (a plain copy, or the int64 -> int32 row conversion with the range check
// the device kernel would do) while the DMA of the buffers before it runs on
// the caller's stream; a download runs the other way (DMA into a buffer,
// host threads copy it out).  Rings are taken from a process-wide free list and
// kept until the process ends; host threads come from one process-wide pool
// shared by every caller (session workers upload concurrently).
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "hpdex_host.h"

namespace hpdex_host {

thread_local bool t_plain_copies = false;

namespace {

constexpr size_t kStageChunk = (size_t)16 << 20;   // bytes per pinned buffer
constexpr int kStageDepth = 3;                      // buffers per ring
constexpr size_t kStageMin = (size_t)4 << 20;       // smaller copies go straight

// Host threads that run the parts of a chunk's fill in parallel with the
// calling thread.  A job stays in the queue until its caller removes it, and
// the caller returns only when every part has run and no worker holds the
// job any more.
class CopyPool {
public:
    static CopyPool& get() {
        static CopyPool pool;
        return pool;
    }
    int width() const { return (int)workers_.size() + 1; }
    // f(i) for i in [0, n): the pool's threads and the caller take parts until
    // none is left; returns when every part has run
    void run(int n, const std::function<void(int)>& f) {
        if (n <= 0) return;
        Job job;
        job.f = &f;
        job.n = n;
        job.left.store(n);
        {
            std::lock_guard<std::mutex> lk(mu_);
            jobs_.push_back(&job);
        }
        cv_.notify_all();
        work(job);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return job.left.load() == 0 && job.users == 0; });
        jobs_.erase(std::find(jobs_.begin(), jobs_.end(), &job));
    }

private:
    struct Job {
        const std::function<void(int)>* f = nullptr;
        int n = 0;
        std::atomic<int> next{0};
        std::atomic<int> left{0};
        int users = 0;   // pool threads inside work() on this job (guarded by mu_)
    };
    CopyPool() {
        const unsigned hc = std::thread::hardware_concurrency();
        const int nw = (int)std::min<unsigned>(7u, hc > 2u ? hc / 2u : 1u);
        for (int i = 0; i < nw; ++i) workers_.emplace_back([this] { loop(); });
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (std::thread& t : workers_) t.join();
    }
    // run the parts of `job` this thread can still take
    void work(Job& job) {
        for (;;) {
            const int i = job.next.fetch_add(1);
            if (i >= job.n) return;
            (*job.f)(i);
            if (job.left.fetch_sub(1) == 1) {
                std::lock_guard<std::mutex> lk(mu_);
                done_.notify_all();
            }
        }
    }
    Job* open_job() {   // a queued job with parts left (mu_ held)
        for (Job* j : jobs_)
            if (j->next.load() < j->n) return j;
        return nullptr;
    }
    void loop() {
        for (;;) {
            Job* job = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu_);
                // the job found open is the one taken: parts are handed out
                // outside the lock, so a second search could find none
                cv_.wait(lk, [&] { return stop_ || (job = open_job()) != nullptr; });
                if (stop_) return;
                ++job->users;
            }
            work(*job);
            {
                std::lock_guard<std::mutex> lk(mu_);
                --job->users;
                done_.notify_all();
            }
        }
    }
    std::vector<std::thread> workers_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    std::deque<Job*> jobs_;
    bool stop_ = false;
};

struct Ring {
    void* buf[kStageDepth] = {};
    hipEvent_t ev[kStageDepth] = {};
};

std::mutex g_ring_mu;
std::vector<Ring*> g_free_rings;

Ring* acquire_ring() {
    {
        std::lock_guard<std::mutex> lk(g_ring_mu);
        if (!g_free_rings.empty()) {
            Ring* r = g_free_rings.back();
            g_free_rings.pop_back();
            return r;
        }
    }
    Ring* r = new Ring();
    for (int i = 0; i < kStageDepth; ++i) {
        if (hipHostMalloc(&r->buf[i], kStageChunk, hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&r->ev[i], hipEventDisableTiming) != hipSuccess) {
            for (int j = 0; j <= i; ++j) {
                if (r->buf[j]) (void)hipHostFree(r->buf[j]);
                if (r->ev[j]) (void)hipEventDestroy(r->ev[j]);
            }
            delete r;
            return nullptr;
        }
    }
    return r;
}

void release_ring(Ring* r) {
    std::lock_guard<std::mutex> lk(g_ring_mu);
    g_free_rings.push_back(r);
}

// Copy `bytes` of device memory at dst from the staging fills: fill(stage,
// off, n) writes bytes [off, off + n) of the source image into `stage`
// (called from several threads on disjoint ranges).  Returns after the last
// DMA has completed, so the ring is free and the caller's memory may change.
hipError_t staged(void* dst, size_t bytes, hipStream_t s,
                  const std::function<void(char* stage, size_t off, size_t n)>& fill,
                  size_t align) {
    Ring* ring = acquire_ring();
    if (!ring) return hipErrorOutOfMemory;
    CopyPool& pool = CopyPool::get();
    const int parts = pool.width();
    hipError_t e = hipSuccess;
    bool used[kStageDepth] = {};
    size_t k = 0;
    for (size_t off = 0; off < bytes && e == hipSuccess; off += kStageChunk, ++k) {
        const int slot = (int)(k % kStageDepth);
        if (used[slot] && (e = hipEventSynchronize(ring->ev[slot])) != hipSuccess) break;
        const size_t n = std::min(kStageChunk, bytes - off);
        char* stage = (char*)ring->buf[slot];
        // parts of whole `align`-byte elements
        const size_t elems = n / align, per = (elems + (size_t)parts - 1) / (size_t)parts;
        pool.run(parts, [&](int i) {
            const size_t a = std::min(elems, (size_t)i * per), b = std::min(elems, a + per);
            if (b > a) fill(stage + a * align, off + a * align, (b - a) * align);
        });
        if ((e = hipMemcpyAsync((char*)dst + off, stage, n, hipMemcpyHostToDevice, s)) != hipSuccess)
            break;
        if ((e = hipEventRecord(ring->ev[slot], s)) != hipSuccess) break;
        used[slot] = true;
    }
    for (int i = 0; i < kStageDepth; ++i)
        if (used[i]) {
            const hipError_t w = hipEventSynchronize(ring->ev[i]);
            if (e == hipSuccess) e = w;
        }
    release_ring(ring);
    return e;
}

}  // namespace

hipError_t download_bytes(void* dst, const void* src, size_t bytes, hipStream_t s) {
    auto plain = [&]() {
        const hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
        return e != hipSuccess ? e : hipStreamSynchronize(s);
    };
    if (t_plain_copies || bytes < kStageMin) return plain();
    Ring* ring = acquire_ring();
    if (!ring) return plain();   // no pinned memory to stage through
    CopyPool& pool = CopyPool::get();
    const int parts = pool.width();
    const size_t nchunks = (bytes + kStageChunk - 1) / kStageChunk;
    hipError_t e = hipSuccess;
    // chunk k: its DMA lands in buffer k % kStageDepth, then host threads copy
    // it out; a buffer is copied out before the chunk kStageDepth later lands
    auto drain = [&](size_t k) {
        const int slot = (int)(k % kStageDepth);
        const hipError_t w = hipEventSynchronize(ring->ev[slot]);
        if (w != hipSuccess) {
            if (e == hipSuccess) e = w;
            return;
        }
        const size_t off = k * kStageChunk, n = std::min(kStageChunk, bytes - off);
        const char* stage = (const char*)ring->buf[slot];
        const size_t per = (n + (size_t)parts - 1) / (size_t)parts;
        pool.run(parts, [&](int i) {
            const size_t a = std::min(n, (size_t)i * per), b = std::min(n, a + per);
            if (b > a) std::memcpy((char*)dst + off + a, stage + a, b - a);
        });
    };
    size_t issued = 0;
    for (size_t k = 0; k < nchunks && e == hipSuccess; ++k) {
        const int slot = (int)(k % kStageDepth);
        if (k >= (size_t)kStageDepth) drain(k - kStageDepth);
        if (e != hipSuccess) break;
        const size_t off = k * kStageChunk, n = std::min(kStageChunk, bytes - off);
        if ((e = hipMemcpyAsync(ring->buf[slot], (const char*)src + off, n, hipMemcpyDeviceToHost,
                                s)) != hipSuccess)
            break;
        if ((e = hipEventRecord(ring->ev[slot], s)) != hipSuccess) break;
        issued = k + 1;
    }
    // the chunks still in the buffers (after an error: only waited for)
    for (size_t k = issued > (size_t)kStageDepth ? issued - kStageDepth : 0; k < issued; ++k) {
        if (e == hipSuccess) drain(k);
        else (void)hipEventSynchronize(ring->ev[k % kStageDepth]);
    }
    release_ring(ring);
    return e;
}

hipError_t upload_bytes(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (t_plain_copies || bytes < kStageMin)
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
    const hipError_t e = staged(dst, bytes, s, [&](char* stage, size_t off, size_t n) {
        std::memcpy(stage, (const char*)src + off, n);
    }, 1);
    if (e == hipErrorOutOfMemory)   // no pinned memory to stage through
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
    return e;
}

hipError_t upload_rows_i64(int32_t* dst, const int64_t* src, size_t n, int64_t R, hipStream_t s) {
    if (t_plain_copies || n * 4 < kStageMin) return hipErrorNotSupported;
    return staged(dst, n * 4, s, [&](char* stage, size_t off, size_t bytes) {
        int32_t* out = (int32_t*)stage;
        const int64_t* in = src + off / 4;
        const size_t m = bytes / 4;
        for (size_t i = 0; i < m; ++i) {
            const int64_t r = in[i];
            out[i] = (r >= 0 && r < R) ? (int32_t)r : -1;
        }
    }, 4);
}

}  // namespace hpdex_host
