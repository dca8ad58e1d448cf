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
#include <memory>
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
        // at most 15 (a GPU's share of a host's cores); the compact uploads'
        // fills take all of them
        const int nw = (int)std::min<unsigned>(15u, hc > 2u ? hc - 1u : 1u);
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
// With `abort`: a fill may set it to stop the copy (no later chunk is filled
// or sent); staged then returns hipErrorNotSupported once the chunks already
// sent have landed.
hipError_t staged(void* dst, size_t bytes, hipStream_t s,
                  const std::function<void(char* stage, size_t off, size_t n)>& fill,
                  size_t align, const std::atomic<bool>* abort = nullptr) {
    Ring* ring = acquire_ring();
    if (!ring) return hipErrorOutOfMemory;
    CopyPool& pool = CopyPool::get();
    const int parts = pool.width();
    hipError_t e = hipSuccess;
    bool used[kStageDepth] = {};
    size_t k = 0;
    for (size_t off = 0; off < bytes && e == hipSuccess; off += kStageChunk, ++k) {
        if (abort && abort->load()) {
            e = hipErrorNotSupported;
            break;
        }
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
        if (abort && abort->load()) {
            e = hipErrorNotSupported;
            break;
        }
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
    // copies into the caller's memory fault its fresh pages in: more than 8
    // threads at once slowed them (38 MB of results: 2.0 -> 8.5 ms with 16)
    const int parts = std::min(pool.width(), 8);
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

namespace {

// Compact words (row & 0xFFFF) << 16 | value of entries a .. b-1 of one
// column, every one of which has a predecessor in the column (entry a - 1);
// returns 0 when one of them does not qualify: its row outside [0, R), not
// above its predecessor's or 2^16 or more above it, or its value not a whole
// number in [0, 65535] with the sign bit clear (-0.0 and NaN do not
// qualify).  Written without branches so that it vectorizes: the checks of a
// segment are or-ed together and tested once.
template <typename IDX>
static inline uint32_t compact_run(uint32_t* __restrict out, const float* __restrict vals,
                                   const IDX* __restrict rows, int64_t a, int64_t b, int64_t R) {
    uint32_t bad = 0u;
    const int64_t n = b - a;
    const IDX* rp = rows + a;
    const float* vp = vals + a;
#pragma clang loop vectorize(enable) interleave(enable)
    for (int64_t i = 0; i < n; ++i) {
        const int64_t r = (int64_t)rp[i], pr = (int64_t)rp[i - 1];
        const float v = vp[i];
        const float vc = v > 0.0f ? (v < 65535.0f ? v : 65535.0f) : 0.0f;
        const int32_t u = (int32_t)vc;
        bad |= (uint32_t)((uint64_t)(r - pr - 1) >= 65535u) |
               (uint32_t)((uint64_t)r >= (uint64_t)R) | (uint32_t)((float)u != v) |
               (__builtin_bit_cast(uint32_t, v) >> 31);
        out[i] = ((uint32_t)r << 16) | (uint32_t)u;
    }
    return bad ? 0u : 1u;
}

// Entries [e0, e0 + m) of a CSC matrix as compact words.  A column with an
// entry that does not qualify (for its first entry: its row in [0, R) and its
// value) is reported through bad(c, entries of the column in the range); its
// words are left as they come.
template <typename IDX, uint32_t (*RUN)(uint32_t*, const float*, const IDX*, int64_t, int64_t,
                                        int64_t), typename BAD>
static void compact_fill_with(uint32_t* out, const float* vals, const IDX* rows,
                              const int64_t* indptr, size_t C, size_t e0, size_t m, int64_t R,
                              BAD&& bad) {
    size_t c = (size_t)(std::upper_bound(indptr, indptr + C + 1, (int64_t)e0) - indptr) - 1;
    int64_t e = (int64_t)e0;
    const int64_t end = (int64_t)(e0 + m);
    while (e < end) {
        const int64_t b = std::min(end, indptr[c + 1]);
        const int64_t s0 = e;
        uint32_t ok = 1u;
        if (e == indptr[c] && e < b) {   // the column's first entry
            const int64_t r = (int64_t)rows[e];
            const float v = vals[e];
            const float vc = v > 0.0f ? (v < 65535.0f ? v : 65535.0f) : 0.0f;
            const int32_t u = (int32_t)vc;
            ok &= (uint32_t)((uint64_t)r < (uint64_t)R) & (uint32_t)((float)u == v) &
                  (uint32_t)((__builtin_bit_cast(uint32_t, v) >> 31) == 0u);
            out[e - (int64_t)e0] = ((uint32_t)r << 16) | (uint32_t)u;
            ++e;
        }
        if (e < b) ok &= RUN(out + (e - (int64_t)e0), vals, rows, e, b, R);
        if (!ok) bad(c, (size_t)(b - s0));
        e = b;
        ++c;
    }
}

template <typename IDX>
static uint32_t compact_run_base(uint32_t* out, const float* vals, const IDX* rows, int64_t a,
                                 int64_t b, int64_t R) {
    return compact_run<IDX>(out, vals, rows, a, b, R);
}
template <typename IDX>
__attribute__((target("avx2"))) static uint32_t compact_run_avx2(uint32_t* out, const float* vals,
                                                                 const IDX* rows, int64_t a,
                                                                 int64_t b, int64_t R) {
    return compact_run<IDX>(out, vals, rows, a, b, R);
}

template <typename IDX, typename BAD>
void compact_fill(uint32_t* out, const float* vals, const IDX* rows, const int64_t* indptr,
                  size_t C, size_t e0, size_t m, int64_t R, BAD&& bad) {
    static const bool avx2 = __builtin_cpu_supports("avx2");
    if (avx2)
        compact_fill_with<IDX, compact_run_avx2<IDX>>(out, vals, rows, indptr, C, e0, m, R, bad);
    else
        compact_fill_with<IDX, compact_run_base<IDX>>(out, vals, rows, indptr, C, e0, m, R, bad);
}

}  // namespace

template <typename IDX>
hipError_t upload_compact(uint32_t* words, int32_t* hi_dev, const float* vals, const IDX* rows,
                          const int64_t* indptr, size_t C, size_t nnz, int64_t R, hipStream_t s,
                          std::vector<int32_t>& patch) {
    patch.clear();
    if (t_plain_copies || nnz * 4 < kStageMin) return hipErrorNotSupported;
    // the high bits of each column's first row (a row out of range makes the
    // column one to patch)
    std::vector<int32_t> hi(C, 0);
    for (size_t c = 0; c < C; ++c)
        if (indptr[c] < indptr[c + 1]) {
            const int64_t r = (int64_t)rows[indptr[c]];
            hi[c] = (r >= 0 && r < R) ? (int32_t)(r >> 16) : 0;
        }
    // columns with an entry that does not qualify; more than an eighth of the
    // entries in them ends the compact copy
    std::unique_ptr<std::atomic<uint8_t>[]> bad_col(new std::atomic<uint8_t>[C]);
    for (size_t c = 0; c < C; ++c) bad_col[c].store(0, std::memory_order_relaxed);
    std::atomic<size_t> bad_entries{0};
    std::atomic<bool> stop{false};
    const size_t limit = nnz / 8;
    hipError_t e = staged(words, nnz * 4, s, [&](char* stage, size_t off, size_t n) {
        if (stop.load()) return;
        compact_fill<IDX>((uint32_t*)stage, vals, rows, indptr, C, off / 4, n / 4, R,
                          [&](size_t c, size_t k) {
                              bad_col[c].store(1, std::memory_order_relaxed);
                              if (bad_entries.fetch_add(k) + k > limit) stop.store(true);
                          });
    }, 4, &stop);
    if (e != hipSuccess) return e == hipErrorOutOfMemory ? hipErrorNotSupported : e;
    for (size_t c = 0; c < C; ++c)
        if (bad_col[c].load(std::memory_order_relaxed)) patch.push_back((int32_t)c);
    if ((e = hipMemcpyAsync(hi_dev, hi.data(), C * 4, hipMemcpyHostToDevice, s)) != hipSuccess)
        return e;
    return hipStreamSynchronize(s);   // hi is a host temporary
}

template hipError_t upload_compact<int32_t>(uint32_t*, int32_t*, const float*, const int32_t*,
                                            const int64_t*, size_t, size_t, int64_t, hipStream_t,
                                            std::vector<int32_t>&);
template hipError_t upload_compact<int64_t>(uint32_t*, int32_t*, const float*, const int64_t*,
                                            const int64_t*, size_t, size_t, int64_t, hipStream_t,
                                            std::vector<int32_t>&);

}  // namespace hpdex_host
