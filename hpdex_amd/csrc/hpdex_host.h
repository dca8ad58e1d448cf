// hpdex_host.h -- host-side internals shared by the C-ABI (hpdex_api.cpp) and
// the serving session (hpdex_session.cpp).  Not part of the public C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "hpdex_internal.h"
#include "hpdex_mi355x.h"

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            return hpdex_host::set_err(err, err_len, std::string("HIP error: ") +          \
                                                         hipGetErrorString(e_) + " (" #x ")"); \
    } while (0)

namespace hpdex_host {

using namespace hpdex;

// Largest exact-null DP workspace per distinct (n1, n2), in doubles.
constexpr size_t kExactMaxCells = (size_t)1 << 27;

struct Tuning {
    int hist_max = kHistMaxBins;
    int size_ordered = 1;
    int count_lds = kCountLdsDefault;
    int count_block = kCountBlockDefault;
    int count_packed = 1;
    int split_arrival = 1;
    int ref_lds_keys = kRefKeysF32;
    int ref_cache = 1;
};

extern std::mutex g_mu;
extern Tuning g_tune;
extern bool g_init;
extern int g_cus;
extern int g_max_lds;
extern bool g_count_ready;
extern char g_name[256];

int set_err(char* err, size_t len, const std::string& m);
int ensure_init(char* err, size_t err_len);
Tuning tuning();

// A block of device memory that a job's buffers are carved from (bump
// allocation, freed as a whole when the job ends).  While a session worker
// builds a job, t_block points at the job's block and DevBuf takes its memory
// there instead of calling hipMalloc.
struct BumpBlock {
    char* base = nullptr;
    size_t cap = 0;
    size_t used = 0;
    bool exhausted = false;   // a take() did not fit
    void* take(size_t n) {
        n = (n + 255) & ~(size_t)255;
        if (used + n > cap) {
            exhausted = true;
            return nullptr;
        }
        void* p = base + used;
        used += n;
        return p;
    }
};
extern thread_local BumpBlock* t_block;

// Host-to-device copies of caller memory (hpdex_stage.cpp): large copies go
// through pinned staging buffers filled by host threads.  t_plain_copies
// (set by the calls that may allocate nothing) makes them plain
// hipMemcpyAsync calls; upload_rows_i64 then returns hipErrorNotSupported
// and the caller converts the rows on the device.
extern thread_local bool t_plain_copies;
hipError_t upload_bytes(void* dst, const void* src, size_t bytes, hipStream_t s);
// device to host, returning when the bytes are in dst (the stream's earlier
// work included)
hipError_t download_bytes(void* dst, const void* src, size_t bytes, hipStream_t s);
// int64 row indices to int32 on the way (rows outside [0, R) become -1)
hipError_t upload_rows_i64(int32_t* dst, const int64_t* src, size_t n, int64_t R, hipStream_t s);
// A float32 CSC matrix in 4 bytes per entry instead of 8: one word (row &
// 0xFFFF) << 16 | value per entry into `words` (nnz of them; the matrix's rows
// array, which launch_decode_compact turns into rows and values) and the high
// bits of each column's first row into hi_dev (C int32).  An entry qualifies
// when its value is a whole number in [0, 65535] (sign bit clear) and its row
// lies in [0, R), above the column's previous row by less than 2^16; the
// columns holding one that does not are listed in `patch` (their words are
// not valid: the caller writes their rows and values as they are after the
// decode).  With more than an eighth of the entries in such columns, for
// copies too small to stage, or when copies are plain, it returns
// hipErrorNotSupported and the caller uploads the matrix as it is.  Returns
// when the copies are done.
template <typename IDX>
hipError_t upload_compact(uint32_t* words, int32_t* hi_dev, const float* vals, const IDX* rows,
                          const int64_t* indptr, size_t C, size_t nnz, int64_t R, hipStream_t s,
                          std::vector<int32_t>& patch);

struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    bool owned = false;   // hipMalloc'd (freed here) rather than carved from a block
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { reset(); }
    void reset() {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        n = 0;
        owned = false;
    }
    hipError_t alloc(size_t bytes) {
        reset();
        n = bytes;
        if (t_block) {
            p = t_block->take(bytes ? bytes : 16);
            return p ? hipSuccess : hipErrorOutOfMemory;
        }
        owned = true;
        return hipMalloc(&p, bytes ? bytes : 16);
    }
    template <typename T> T* as() const { return (T*)p; }
};

struct Stream {
    hipStream_t s = nullptr;
    ~Stream() {
        if (s) (void)hipStreamDestroy(s);
    }
    hipError_t create() { return hipStreamCreateWithFlags(&s, hipStreamNonBlocking); }
};

// A resident matrix's reference tables for one set of reference rows
// (RefCacheCol per column, the sorted keys and the bucket indexes), copied
// from a finished job and read by later jobs with the same reference rows.
// Immutable once installed; a job that reads it holds a reference, so a
// replaced entry stays on the device until its last reader ends.  Its
// buffers are the matrix's (hipMalloc'd), not a job's.
struct RefCacheEntry {
    std::vector<uint64_t> ref_rows;   // bit r: row r is in the reference group
    std::vector<uint8_t> valid;       // host copy: column c has tables
    DevBuf cols, keys, tab;
    int64_t n_cols = 0;               // columns with tables
    size_t bytes = 0;
};

// Bitmask of the rows of group 0 (the reference) in R labels.
std::vector<uint64_t> reference_rows(const int32_t* group_id, size_t R);

bool valid_dtype(int d);
size_t dtype_size(int d);
int check_indptr(const int64_t* indptr, size_t C, size_t nnz, char* err, size_t err_len);

}  // namespace hpdex_host

// Resident matrix (the C-ABI's opaque hpdex_matrix).
struct hpdex_matrix {
    int dtype = 0;
    bool f64 = false;
    size_t R = 0, C = 0, nnz = 0;
    hpdex_host::DevBuf vals, rows, indptr;
    hpdex_host::DevBuf colrange;    // 2C int32: value range of integer-valued columns
    hpdex_host::DevBuf colident;    // C bytes: 1 = the column stores rows 0 .. R-1 in order
    hpdex_host::DevBuf colint;      // C bytes: device copy of col_int
    bool any_ident = false;         // some column stores rows 0 .. R-1 in order
    std::vector<int64_t> h_indptr;
    std::vector<int32_t> order;     // columns, largest first
    std::vector<uint8_t> col_int;   // 1: every stored value is an integer in 16 bits (or
                                    // non-finite) -- the column may go to the count kernel
    std::vector<int32_t> h_range;   // host copy of colrange
    bool pack = true;               // make packed entries (a matrix that is queried more
                                    // than once: they cost one pass over the entries)
    hpdex_host::DevBuf packed;      // packed entries (pack_row_bits) of the packable columns
    uint32_t pack_vb = 0;           // their value bits
    std::vector<uint8_t> col_packed;   // 1: column c has packed entries
    std::vector<uint32_t> h_dense;     // packed columns: distinct nonzero values (their bins)
    std::vector<uint32_t> h_zneg;      // packed columns: distinct values below zero
    // the sorted reference of the last sort-path job whose reference rows had
    // no tables yet (resident matrices only: those created through the C-ABI)
    bool resident = true;
    bool compact = false;   // uploaded in 4 bytes per entry (upload_compact)
    int64_t compact_patched = 0;   // of its columns, those written as they were
    mutable std::mutex rc_mu;
    mutable std::shared_ptr<hpdex_host::RefCacheEntry> rc;
};

namespace hpdex_host {

// Per-job device state shared by the one-shot call, the plan and the session.
struct JobState {
    size_t T = 0, G = 0, C = 0;
    bool stats = false;          // tie and sample sizes kept per pair
    hpdex_mwu_opts opts{};
    std::vector<uint32_t> group_rows;
    DevBuf gid, gidp, grows, U1, P, tie, n1, n2, flags;  // flags: err_key, status_col, status
    DevBuf members, member_group;   // the groups' rows, group by group (matrices with
    uint32_t n_members = 0;         // columns in row order)
    uint32_t gid_bits = 0;
    size_t gid_words = 0;
    DevBuf keys, cnt, zcnt, goff, plan, reftab;     // pipeline workspace
    DevBuf spill, spill_off;                        // reference slices beyond the chip
    DevBuf d_job;                                   // descriptor
    DevBuf d_count_items, d_gen_items;              // work queues of the two paths
    std::vector<CountItem> h_count_items;           // host copies (batched launches merge them)
    std::vector<WorkItem> h_gen_items;
    JobDesc J{};                                    // host copy of the descriptor
    int n_count = 0, n_count_cols = 0, n_gen_static = 0;   // count workgroups, their columns
    int n_count_packed = 0;                                  // of them read packed
    int64_t count_nnz = 0;
    uint32_t count_lds = 0;
    int count_block = kCountBlockDefault;
    // exact null tables
    std::vector<ExactTable> tabs;
    std::vector<int32_t> tidx_by_g;
    DevBuf d_tabs, tabbuf, tscratch, d_tidx;
    size_t scratch_per_table = 0;
    bool exact_by_pair = false;
    // the matrix's reference tables this job reads (rc_use), or, when it has
    // none for these reference rows, a fill of them after the job's first run
    std::shared_ptr<RefCacheEntry> rc_use;
    std::vector<uint64_t> ref_rows;
    bool rc_fill = false;
    int64_t rc_cols = 0;   // sort-path columns whose tables the cache holds

    unsigned long long* err_key() const { return flags.as<unsigned long long>(); }
    unsigned long long* status_col() const { return flags.as<unsigned long long>() + 1; }
    uint32_t* status() const { return (uint32_t*)(flags.as<unsigned long long>() + 2); }
};

int finish_matrix(hpdex_matrix* m, hipStream_t s, char* err, size_t err_len);
int widen_indptr(const void* p, int bits, size_t n, size_t nnz, std::vector<int64_t>& out,
                 char* err, size_t err_len);
int upload_matrix_from(hpdex_matrix* m, const hpdex_matrix_src* src, hipStream_t s, char* err,
                       size_t err_len);
int upload_matrix(hpdex_matrix* m, int dtype, const void* data, const int64_t* indices,
                  const int64_t* indptr, size_t R, size_t C, size_t nnz, hipStream_t s,
                  char* err, size_t err_len);
int job_setup(JobState& js, const hpdex_matrix* m, const int32_t* group_id, size_t T,
              const hpdex_mwu_opts* o, bool stats, hipStream_t s, char* err, size_t err_len);
JobDesc make_desc(const JobState& js, const hpdex_matrix* m, const Tuning& tu);
int job_upload_desc(JobState& js, const JobDesc& J, const hpdex_matrix& m, const Tuning& tu,
                    bool size_ordered, hipStream_t s, char* err, size_t err_len);
PipelineLaunch pipeline_of(const JobState& js, bool f64);
int exact_plan(JobState& js, const double* n1pair, const double* n2pair, hipStream_t s,
               char* err, size_t err_len);
int exact_launch(const JobState& js, hipStream_t s, bool build_tables, char* err,
                 size_t err_len);
int job_errors(const JobState& js, hipStream_t s, char* err, size_t err_len);
int copy_out(const JobState& js, double* U1, double* P, double* tie, double* n1, double* n2,
             hipStream_t s, char* err, size_t err_len);
bool opts_ok(const hpdex_mwu_opts* o);
int run_exact_none(JobState& js, hipStream_t s, char* err, size_t err_len);
// After a job's launch sequence has completed on s: when the job was set up
// to fill its matrix's reference cache (js.rc_fill), copy its sort-path
// columns' reference tables into a new entry (device memory of the matrix)
// and install it.  The fill is skipped when the job recorded an error.
int refcache_fill(JobState& js, const hpdex_matrix& m, hipStream_t s, char* err, size_t err_len);

// Device bytes, as BumpBlock::take rounds every DevBuf (256 bytes, at least
// one 256-byte unit), of:
//   matrix_footprint  -- upload_matrix / upload_matrix_from for a CSC (int32 or
//                        int64 indices) or CSR source, staging included, with
//                        packed entries when `pack`;
//   job_footprint     -- job_setup + job_upload_desc + the exact-null tables of
//                        zero_handling min/max/mix (group_rows: G row counts;
//                        may be null when the job is asymptotic).  The tables
//                        of exact p in mode none depend on the data and are
//                        not included;
//   de_footprint      -- everything one hpdex_de_csc call allocates.
// Upper bounds: count-kernel items are counted at kMaxSlices per column.
size_t up256(size_t n);
size_t matrix_footprint(int dtype, size_t R, size_t C, size_t nnz, int layout, int index_bits,
                        bool pack);
size_t job_footprint(size_t R, size_t C, size_t nnz, bool f64, size_t T, const hpdex_mwu_opts& o,
                     const uint32_t* group_rows, bool stats, size_t spill_keys);
// Spill keys of a column of len entries whose reference has gr0 rows (0 when
// its reference slice fits the on-chip sort of cap keys), and their bound over
// a matrix (indptr null: min(nnz, C * gr0), for a matrix not yet in CSC).
size_t spill_keys_of(int64_t len, uint32_t gr0, size_t cap);
size_t spill_keys_bound(const int64_t* indptr, size_t C, size_t nnz, uint32_t gr0, bool f64,
                        const Tuning& tu);
// reference keys a column's slice may hold and still be sorted in LDS
size_t ref_lds_cap(bool f64, const Tuning& tu);
size_t de_footprint(int dtype, size_t R, size_t C, size_t nnz, size_t T);

// Several jobs of one value type as one batched launch: their descriptors,
// every job's count-kernel items (naming the job) and sort-path columns, each
// list ordered largest first.
struct MergedJobs {
    std::vector<JobDesc> jobs;
    std::vector<CountItem> count;
    std::vector<WorkItem> gen;
    uint32_t lds = 0;
    int G = 2;
};
void merge_jobs(const std::vector<const JobState*>& jss, const std::vector<const hpdex_matrix*>& ms,
                MergedJobs& out);
// Put a merged batch on the device (reusing the buffers when they are large
// enough) and describe its launch.
int upload_merged(const MergedJobs& mj, bool f64, const Tuning& tu, DevBuf& d_jobs,
                  DevBuf& d_count, DevBuf& d_gen, hipStream_t s, BatchLaunch& L, char* err,
                  size_t err_len);

int mwu_oneshot(int dtype, const void* data, const int64_t* indices, const int64_t* indptr,
                size_t n_rows, size_t n_cols, size_t nnz, const int32_t* group_id,
                size_t n_targets, const hpdex_mwu_opts* opts, double* U1_out, double* P_out,
                double* tie_out, double* n1_out, double* n2_out, hipStream_t s, char* err,
                size_t err_len);

}  // namespace hpdex_host
