// hpdex_session.cpp -- serving session: DE jobs of different sizes admitted by
// HBM footprint and run concurrently on one GPU (continuous batching: the
// work items of every job ready at the time run in one merged launch).
//
// One arena of device memory is allocated when the session starts.  Each job's
// footprint -- an upper bound of every buffer it allocates (its uploaded matrix
// if it brings one, the per-job workspace of hpdex_api.cpp, exact-null tables)
// -- is computed when it is submitted.  Worker threads, one stream each, take
// jobs from a FIFO queue whenever the head job's footprint fits in the arena;
// the job's buffers are then carved from its block (BumpBlock) with no
// hipMalloc; the worker uploads and sets the job up on its own stream, then
// hands it to the launcher thread, which runs the kernels of every job
// waiting at that moment as one merged launch (one queue of all their
// count-kernel items, largest first, and one list of their sort-path
// columns) while the next jobs upload; the worker then copies the results
// out and the block returns to the arena.  Jobs that name a resident matrix skip the upload:
// a new target set against cells already in HBM costs only its workspace.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <utility>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "hpdex_host.h"
#include "hpdex_internal.h"
#include "hpdex_mi355x.h"

using namespace hpdex;
using namespace hpdex_host;

namespace {

using Clock = std::chrono::steady_clock;

// First-fit allocator over [0, cap) of the arena, with coalescing of free ranges.
class Arena {
  public:
    bool take(size_t n, size_t* off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            const size_t o = it->first, rem = it->second - n;
            free_.erase(it);
            if (rem) free_[o + n] = rem;
            used_ += n;
            *off = o;
            return true;
        }
        return false;
    }
    void give(size_t off, size_t n) {
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
        used_ -= n;
    }
    void reset(size_t cap) {
        free_.clear();
        free_[0] = cap;
        used_ = 0;
    }
    size_t used() const { return used_; }

  private:
    std::map<size_t, size_t> free_;
    size_t used_ = 0;
};

struct JobRec {
    hpdex_job job{};
    size_t footprint = 0;
    Clock::time_point t_submit, t_start, t_end;
    Clock::time_point t_uploaded, t_ready, t_ran;   // phases of the run (see hpdex_job_times)
    bool done = false;
    int rc = 0;
    std::string err;
};

// The host matrix a job brings, as hpdex_matrix_create_from takes it.
hpdex_matrix_src source_of(const hpdex_job& j) {
    hpdex_matrix_src src{};
    src.layout = j.layout;
    src.dtype = j.dtype;
    src.data = j.data;
    src.indices = j.indices;
    src.indptr = j.indptr;
    src.index_bits = j.index_bits == 32 ? 32 : 64;
    src.n_rows = j.n_rows;
    src.n_cols = j.n_cols;
    src.nnz = j.nnz;
    return src;
}

// Upper bound of what one job allocates through DevBuf: its matrix, if it
// brings one, and its workspace (the library's own footprint formulas).
size_t session_job_footprint(const hpdex_job& j, const std::vector<uint32_t>& group_rows) {
    const hpdex_matrix* m = j.matrix;
    const size_t R = m ? m->R : j.n_rows, C = m ? m->C : j.n_cols, nnz = m ? m->nnz : j.nnz;
    const bool f64 = m ? m->f64 : dtype_uses_double(j.dtype);
    size_t f = 0;
    if (!m)
        f += matrix_footprint(j.dtype, R, C, nnz, j.layout, j.index_bits == 32 ? 32 : 64, false);
    // the column pointers bound the spill space: the resident matrix's, or
    // the CSC's the job brings (a CSR job's columns are known only on the device)
    std::vector<int64_t> ip;
    const int64_t* indptr = m ? m->h_indptr.data() : nullptr;
    if (!m && j.layout == HPDEX_LAYOUT_CSC) {
        ip.resize(C + 1);
        for (size_t c = 0; c <= C; ++c)
            ip[c] = j.index_bits == 32 ? (int64_t)((const int32_t*)j.indptr)[c]
                                       : ((const int64_t*)j.indptr)[c];
        indptr = ip.data();
    }
    const size_t spill = spill_keys_bound(indptr, C, nnz, group_rows[0], f64, tuning());
    return f + job_footprint(R, C, nnz, f64, j.n_targets, j.opts, group_rows.data(), false, spill);
}

}  // namespace

struct hpdex_session {
    hpdex_session_opts o{};
    void* base = nullptr;
    Arena arena;
    std::mutex mu;
    std::condition_variable cv_admit;   // workers: a job was queued or memory was freed
    std::condition_variable cv_done;    // waiters: a job ended
    std::deque<std::shared_ptr<JobRec>> pending;
    std::map<int64_t, std::shared_ptr<JobRec>> jobs;
    int64_t next_ticket = 1;
    bool stop = false;
    int in_flight = 0, peak_in_flight = 0;
    size_t peak_bytes = 0;
    std::vector<std::thread> workers;
    // continuous batching: jobs set up and waiting for their kernels, and the
    // launcher that runs every waiting job's work items as one merged launch
    std::deque<struct ReadyJob*> ready;
    std::condition_variable cv_ready;   // launcher: a job is ready (or the session stops)
    std::condition_variable cv_ran;     // workers: a batch holding their job ended
    std::thread launcher;
    bool stop_launcher = false;
    int64_t batches = 0, batched_jobs = 0, max_batch = 0;
};

namespace {

}  // namespace

// A job whose matrix, labels and work lists are on the device, waiting for
// the launcher to run its kernels in the next merged launch.
struct ReadyJob {
    const JobState* js;
    const hpdex_matrix* m;
    bool ran = false;
    int rc = 0;
    std::string err;
};

namespace {

// Everything one job does, on the worker's stream, with buffers from its
// block: the upload (when it brings its matrix), labels, routing and exact
// tables; then its kernels, in a launch the launcher merges with every other
// job ready at the time; then the exact p lookup, the errors and the copy of
// its results to the caller's buffers.
void run_job(hpdex_session* S, JobRec& R, hipStream_t s, char* block, size_t cap) {
    BumpBlock blk;
    blk.base = block;
    blk.cap = cap;
    t_block = &blk;
    char err[1024];
    const size_t err_len = sizeof err;
    err[0] = 0;
    int rc = 0;
    const hpdex_job& j = R.job;
    const Tuning tu = tuning();
    {
        std::unique_ptr<hpdex_matrix> own;
        const hpdex_matrix* m = j.matrix;
        JobState js;
        do {
            if (!m) {
                own.reset(new hpdex_matrix());
                own->pack = false;   // the job's own matrix serves this job only
                own->resident = false;
                const hpdex_matrix_src src = source_of(j);
                rc = upload_matrix_from(own.get(), &src, s, err, err_len);
                if (rc) break;
                m = own.get();
            }
            R.t_uploaded = R.t_ready = R.t_ran = Clock::now();
            if (m->C == 0) break;
            rc = job_setup(js, m, j.group_id, j.n_targets, &j.opts, false, s, err, err_len);
            if (rc) break;
            const JobDesc J = make_desc(js, m, tu);
            rc = job_upload_desc(js, J, *m, tu, tu.size_ordered != 0, s, err, err_len);
            if (rc) break;
            const bool exact = j.opts.method == 1;
            if (exact) {
                rc = exact_plan(js, nullptr, nullptr, s, err, err_len);
                if (rc) break;
            }
            // everything the kernels read is on the device before they run on
            // the launcher's stream
            if (hipStreamSynchronize(s) != hipSuccess) {
                rc = set_err(err, err_len, "HIP error before the job's kernels");
                break;
            }
            ReadyJob rj;
            rj.js = &js;
            rj.m = m;
            R.t_ready = Clock::now();
            {
                std::unique_lock<std::mutex> lk(S->mu);
                S->ready.push_back(&rj);
                S->cv_ready.notify_one();
                S->cv_ran.wait(lk, [&] { return rj.ran; });
            }
            R.t_ran = Clock::now();
            if (rj.rc) {
                rc = set_err(err, err_len, rj.err);
                break;
            }
            if (exact) {
                rc = exact_launch(js, s, true, err, err_len);
                if (rc) break;
            }
            rc = job_errors(js, s, err, err_len);
            if (rc) break;
            rc = copy_out(js, j.U1_out, j.P_out, nullptr, nullptr, nullptr, s, err, err_len);
            if (rc) break;
            // a resident matrix keeps this job's sorted reference for later
            // jobs with the same reference rows
            rc = refcache_fill(js, *m, s, err, err_len);
        } while (0);
        // nothing of this job may still run when its block goes back to the arena
        const hipError_t se = hipStreamSynchronize(s);
        if (!rc && se != hipSuccess)
            rc = set_err(err, err_len, std::string("HIP error: ") + hipGetErrorString(se));
    }
    t_block = nullptr;
    R.rc = rc;
    if (rc) R.err = err;
}

// Pinned host staging of one merged launch's lists (allocated with
// hipHostMalloc, grown as needed), so that their copies to the device are
// asynchronous and the launcher can build the next launch meanwhile.
struct PinnedBuf {
    void* p = nullptr;
    size_t n = 0;
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
    hipError_t fit(size_t bytes) {
        if (p && n >= bytes) return hipSuccess;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
        const hipError_t e = hipHostMalloc(&p, bytes ? bytes : 16, hipHostMallocDefault);
        if (e == hipSuccess) n = bytes ? bytes : 16;
        return e;
    }
};

// One merged launch in flight: its jobs, the event after its kernels, and the
// device and pinned buffers its lists live in until it ends.
struct InFlight {
    std::vector<ReadyJob*> jobs;
    hipEvent_t done = nullptr;
    int rc = 0;
    std::string err;
};
struct LaunchSlot {
    DevBuf d_jobs[2], d_count[2], d_gen[2];
    PinnedBuf h_jobs[2], h_count[2], h_gen[2];
};

// Copy one merged list to the device through its pinned staging, on s.
template <typename T>
hipError_t stage(const std::vector<T>& v, PinnedBuf& h, DevBuf& d, hipStream_t s) {
    const size_t bytes = v.size() * sizeof(T);
    hipError_t e = h.fit(bytes);
    if (e != hipSuccess) return e;
    if (!(d.p && d.n >= bytes) && (e = d.alloc(bytes ? bytes : 16)) != hipSuccess) return e;
    if (!bytes) return hipSuccess;
    std::memcpy(h.p, v.data(), bytes);
    return hipMemcpyAsync(d.p, h.p, bytes, hipMemcpyHostToDevice, s);
}

// The launcher: whenever jobs are ready, take all of them and run their work
// items as one merged launch per value type (the count kernel over one queue
// of every job's items, largest first, then the sort-path kernels over every
// job's other columns).  Launches go to one stream back to back; two may be
// in flight, each with its own buffers, so the next one is built and its
// lists copied while the current one runs; a launch's jobs are released when
// its event has passed.
void launcher(hpdex_session* S) {
    Stream st;
    const hipError_t se = st.create();
    constexpr int kSlots = 2;
    LaunchSlot slots[kSlots];
    std::deque<std::pair<InFlight, int>> flight;   // with its slot
    int next_slot = 0;
    auto release = [&](InFlight& f) {
        const hipError_t e = hipEventSynchronize(f.done);
        if (!f.rc && e != hipSuccess) {
            f.rc = 1;
            f.err = std::string("HIP error: ") + hipGetErrorString(e);
        }
        (void)hipEventDestroy(f.done);
        {
            std::lock_guard<std::mutex> lk(S->mu);
            for (ReadyJob* r : f.jobs) {
                r->rc = f.rc;
                if (f.rc) r->err = f.err;
                r->ran = true;
            }
            S->batches += 1;
            S->batched_jobs += (int64_t)f.jobs.size();
            S->max_batch = std::max(S->max_batch, (int64_t)f.jobs.size());
        }
        S->cv_ran.notify_all();
    };
    for (;;) {
        std::vector<ReadyJob*> batch;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            if (flight.empty())
                S->cv_ready.wait(lk, [&] { return !S->ready.empty() || S->stop_launcher; });
            if (S->ready.empty() && flight.empty()) return;   // stopping, nothing left
            if (!S->ready.empty() && (int)flight.size() < kSlots) {
                batch.assign(S->ready.begin(), S->ready.end());
                S->ready.clear();
            }
        }
        if (batch.empty()) {   // nothing new to launch: finish the oldest launch
            release(flight.front().first);
            flight.pop_front();
            continue;
        }
        LaunchSlot& sl = slots[next_slot];
        InFlight f;
        f.jobs = std::move(batch);
        char err[1024];
        const size_t err_len = sizeof err;
        err[0] = 0;
        int rc = se != hipSuccess ? set_err(err, err_len, "HIP error: launcher stream") : 0;
        const Tuning tu = tuning();
        for (int v = 0; v < 2 && !rc; ++v) {
            std::vector<const JobState*> jss;
            std::vector<const hpdex_matrix*> ms;
            for (ReadyJob* r : f.jobs)
                if ((int)r->m->f64 == v) {
                    jss.push_back(r->js);
                    ms.push_back(r->m);
                }
            if (jss.empty()) continue;
            MergedJobs mj;
            merge_jobs(jss, ms, mj);
            hipError_t e = stage(mj.jobs, sl.h_jobs[v], sl.d_jobs[v], st.s);
            if (e == hipSuccess) e = stage(mj.count, sl.h_count[v], sl.d_count[v], st.s);
            if (e == hipSuccess) e = stage(mj.gen, sl.h_gen[v], sl.d_gen[v], st.s);
            BatchLaunch L{};
            L.jobs = sl.d_jobs[v].as<JobDesc>();
            L.count_items = sl.d_count[v].as<CountItem>();
            L.n_count = (int)mj.count.size();
            L.gen_items = sl.d_gen[v].as<WorkItem>();
            L.n_gen = (int)mj.gen.size();
            L.G = mj.G;
            L.f64 = v == 1;
            L.count_lds = mj.lds ? mj.lds : (uint32_t)tu.count_lds;
            L.count_block = tu.count_block;
            L.split_lds_max = g_count_ready ? (uint32_t)kSplitLdsMax : (uint32_t)kLdsBudget;
            if (e == hipSuccess) e = launch_mwu_batch(L, st.s);
            if (e != hipSuccess)
                rc = set_err(err, err_len, std::string("HIP error: ") + hipGetErrorString(e));
        }
        f.rc = rc;
        if (rc) f.err = err;
        if (hipEventCreateWithFlags(&f.done, hipEventDisableTiming) != hipSuccess ||
            hipEventRecord(f.done, st.s) != hipSuccess) {
            (void)hipStreamSynchronize(st.s);
            if (!f.rc) {
                f.rc = 1;
                f.err = "HIP error: launcher event";
            }
            if (f.done) (void)hipEventDestroy(f.done);
            (void)hipEventCreateWithFlags(&f.done, hipEventDisableTiming);   // passes at once
        }
        flight.emplace_back(std::move(f), next_slot);
        next_slot = (next_slot + 1) % kSlots;
        // a full pipeline: the oldest launch must end before its slot is reused
        if ((int)flight.size() == kSlots) {
            release(flight.front().first);
            flight.pop_front();
        }
    }
}

void worker(hpdex_session* S) {
    Stream st;
    const hipError_t se = st.create();
    for (;;) {
        std::shared_ptr<JobRec> J;
        size_t off = 0;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            for (;;) {
                if (!S->pending.empty()) {
                    if (S->arena.take(S->pending.front()->footprint, &off)) {
                        J = S->pending.front();
                        S->pending.pop_front();
                        break;
                    }
                } else if (S->stop) {
                    return;
                }
                S->cv_admit.wait(lk);
            }
            ++S->in_flight;
            S->peak_in_flight = std::max(S->peak_in_flight, S->in_flight);
            S->peak_bytes = std::max(S->peak_bytes, S->arena.used());
            J->t_start = Clock::now();
            J->t_uploaded = J->t_ready = J->t_ran = J->t_start;
        }
        if (se != hipSuccess) {
            J->rc = 1;
            J->err = std::string("HIP error: ") + hipGetErrorString(se) + " (stream creation)";
        } else {
            run_job(S, *J, st.s, (char*)S->base + off, J->footprint);
        }
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->arena.give(off, J->footprint);
            --S->in_flight;
            J->t_end = Clock::now();
            J->done = true;
        }
        S->cv_admit.notify_all();
        S->cv_done.notify_all();
    }
}

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace

extern "C" {

int hpdex_session_create(const hpdex_session_opts* o, hpdex_session** out, char* err,
                         size_t err_len) {
    if (!o || !out) return set_err(err, err_len, "null argument");
    *out = nullptr;
    if (o->hbm_budget_bytes < ((size_t)1 << 20))
        return set_err(err, err_len, "hbm_budget_bytes must be at least 1 MiB");
    if (o->slots < 1 || o->slots > 64) return set_err(err, err_len, "slots must be in [1, 64]");
    if (ensure_init(err, err_len)) return 1;
    std::unique_ptr<hpdex_session> S(new hpdex_session());
    S->o = *o;
    HIPCHK(hipMalloc(&S->base, o->hbm_budget_bytes));
    S->arena.reset(o->hbm_budget_bytes);
    S->launcher = std::thread(launcher, S.get());
    for (int i = 0; i < o->slots; ++i) S->workers.emplace_back(worker, S.get());
    *out = S.release();
    return 0;
}

int hpdex_session_submit(hpdex_session* S, const hpdex_job* job, int64_t* ticket, char* err,
                         size_t err_len) {
    if (!S || !job || !ticket) return set_err(err, err_len, "null argument");
    const hpdex_job& j = *job;
    if (j.n_targets < 1)
        return set_err(err, err_len, "[mannwhitney] n_targets must be >= 1 (>=1 target groups)");
    if (!opts_ok(&j.opts)) return set_err(err, err_len, "invalid option block");
    if (j.opts.method == 1 && j.opts.zero_handling == 0)
        return set_err(err, err_len,
                       "exact p with zero_handling none sizes its tables from the data: use "
                       "hpdex_mwu_csc for such jobs");
    size_t R, C;
    if (j.matrix) {
        R = j.matrix->R;
        C = j.matrix->C;
    } else {
        if (!valid_dtype(j.dtype)) return set_err(err, err_len, "Unsupported dtype for data");
        if (j.layout != HPDEX_LAYOUT_CSC && j.layout != HPDEX_LAYOUT_CSR)
            return set_err(err, err_len, "a session job brings a CSC or CSR matrix");
        if (j.index_bits != 0 && j.index_bits != 32 && j.index_bits != 64)
            return set_err(err, err_len, "index_bits must be 32 or 64");
        if (!j.indptr || (j.nnz && (!j.data || !j.indices)))
            return set_err(err, err_len, "null matrix array");
        std::vector<int64_t> ptr;
        const bool csr = j.layout == HPDEX_LAYOUT_CSR;
        if (widen_indptr(j.indptr, j.index_bits == 32 ? 32 : 64, csr ? j.n_rows : j.n_cols, j.nnz,
                         ptr, err, err_len))
            return 1;
        if (csr && (ptr[0] != 0 || (size_t)ptr[j.n_rows] != j.nnz))
            return set_err(err, err_len, "CSR indptr must start at 0 and end at nnz");
        R = j.n_rows;
        C = j.n_cols;
    }
    if (R && !j.group_id) return set_err(err, err_len, "null group_id");
    if (C && (!j.U1_out || !j.P_out)) return set_err(err, err_len, "null output buffer");
    std::vector<uint32_t> group_rows(j.n_targets + 1, 0u);
    for (size_t r = 0; r < R; ++r) {
        const int g = j.group_id[r];
        if (g >= 0 && (size_t)g <= j.n_targets) ++group_rows[g];
    }
    auto rec = std::make_shared<JobRec>();
    rec->job = j;
    rec->footprint = session_job_footprint(j, group_rows);
    if (rec->footprint > S->o.hbm_budget_bytes)
        return set_err(err, err_len,
                       "job needs " + std::to_string(rec->footprint) +
                           " bytes of HBM, more than the session's budget of " +
                           std::to_string(S->o.hbm_budget_bytes));
    rec->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) return set_err(err, err_len, "session is closing");
        *ticket = S->next_ticket++;
        S->jobs[*ticket] = rec;
        S->pending.push_back(rec);
    }
    S->cv_admit.notify_all();
    return 0;
}

int hpdex_session_wait(hpdex_session* S, int64_t ticket, hpdex_job_times* times, char* err,
                       size_t err_len) {
    if (!S) return set_err(err, err_len, "null session");
    std::shared_ptr<JobRec> J;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        auto it = S->jobs.find(ticket);
        if (it == S->jobs.end()) return set_err(err, err_len, "unknown ticket");
        J = it->second;
        S->cv_done.wait(lk, [&] { return J->done; });
        S->jobs.erase(ticket);
    }
    if (times) {
        times->queue_ms = ms_between(J->t_submit, J->t_start);
        times->run_ms = ms_between(J->t_start, J->t_end);
        times->latency_ms = ms_between(J->t_submit, J->t_end);
        times->footprint_bytes = J->footprint;
        times->upload_ms = ms_between(J->t_start, J->t_uploaded);
        times->setup_ms = ms_between(J->t_uploaded, J->t_ready);
        times->kernels_ms = ms_between(J->t_ready, J->t_ran);
        times->finish_ms = ms_between(J->t_ran, J->t_end);
    }
    if (J->rc) return set_err(err, err_len, J->err);
    return 0;
}

int hpdex_session_peak(hpdex_session* S, int32_t* jobs_in_flight, size_t* arena_bytes) {
    if (!S) return 1;
    std::lock_guard<std::mutex> lk(S->mu);
    if (jobs_in_flight) *jobs_in_flight = S->peak_in_flight;
    if (arena_bytes) *arena_bytes = S->peak_bytes;
    return 0;
}

int hpdex_session_batches(hpdex_session* S, int64_t* launches, int64_t* jobs, int64_t* max_jobs) {
    if (!S) return 1;
    std::lock_guard<std::mutex> lk(S->mu);
    if (launches) *launches = S->batches;
    if (jobs) *jobs = S->batched_jobs;
    if (max_jobs) *max_jobs = S->max_batch;
    return 0;
}

void hpdex_session_destroy(hpdex_session* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_admit.notify_all();
    for (auto& t : S->workers) t.join();   // their last jobs still need the launcher
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop_launcher = true;
    }
    S->cv_ready.notify_all();
    S->launcher.join();
    if (S->base) (void)hipFree(S->base);
    delete S;
}

}  // extern "C"
