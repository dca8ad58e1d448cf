// hpdex_internal.h -- structures shared by the host scheduler and the gfx950
// kernels.  Not part of the public C-ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace hpdex {

enum ZeroMode : int32_t { kZeroNone = 0, kZeroMin = 1, kZeroMax = 2, kZeroMix = 3 };

// Status bits a kernel raises when a column does not fit its on-chip plan.
enum : uint32_t {
    kStatusCapacity = 1u,   // a reference slice beyond the on-chip sort had no spill space
    kStatusGroups = 2u,     // too many groups for the LDS counters
    kStatusRange = 4u,      // a count-path column held a value outside its recorded range
};

// Launch geometry shared by the per-column kernels.
constexpr int kColBlock = 256;              // threads per workgroup (4 waves)
constexpr int kTargetBlock = 512;           // threads per target-window workgroup (8 waves)
constexpr int kRefBlock = 512;              // threads per reference-kernel workgroup (8 waves)
// Split kernel counters: per group a 64-bit word (explicit zeros in the high
// half, nonzero values in the low) and a 32-bit offset, plus kSplitDummy
// words that entries outside every group hit (one per lane of a half-wave,
// so those no-op atomics meet in no bank).
constexpr uint32_t kSplitDummy = 32;
__host__ __device__ constexpr uint64_t split_counter_bytes(uint64_t G) {
    return (12u * (G + kSplitDummy) + 15u) & ~(uint64_t)15u;
}
constexpr int kSplitBlock = 1024;           // threads per split workgroup (16 waves)
constexpr int kSplitLdsMax = 152 * 1024;    // dynamic LDS a split workgroup may take
constexpr int kLdsBudget = 64 * 1024;       // dynamic LDS per workgroup, at most
constexpr int kHistMaxBins = 8192;          // histogram path: widest integer range
constexpr int kMaxBuckets = 4096;           // sort path: bucket index entries (pow2)
constexpr int kRefKeysF32 = 8192;           // sort path: reference keys kept on chip (4 B keys)
constexpr int kRefKeysF64 = 4096;           // sort path: reference keys kept on chip (8 B keys)
constexpr int kTabStride = (kHistMaxBins > kMaxBuckets ? kHistMaxBins : kMaxBuckets) + 1;

// LDS of the reference-table and target kernels: the larger of a histogram
// prefix (kHistMaxBins + 1 counts) and the on-chip sorted reference with its
// bucket index, rounded to 16 bytes.
__host__ __device__ constexpr uint32_t ref_table_lds(bool f64) {
    const uint32_t hist = 4u * (uint32_t)(kHistMaxBins + 1);
    const uint32_t keys = f64 ? 8u * (uint32_t)kRefKeysF64 : 4u * (uint32_t)kRefKeysF32;
    const uint32_t sort = keys + 4u * (uint32_t)(kMaxBuckets + 1);
    return ((hist > sort ? hist : sort) + 15u) & ~15u;
}

// Target window kernel: per wave, a table of kBigTieSlots (key, target,
// count) words that finds a key's position among the equal keys of its
// target (a window's at most 64 keys fill a quarter of it), followed by the
// wave's 64-byte start map (kStartMapWords words).  A target of 65 to
// kHashTargetKeys 4-byte keys is ranked by the window kernel in chunks of 64
// keys with its tie positions from the same table; larger ones, and 8-byte
// keys, are sorted (the sorted pass).
constexpr uint64_t kTieEmpty = ~0ull;
constexpr uint32_t kTieCountBits = 26;
constexpr uint64_t kTieCountMask = (1ull << kTieCountBits) - 1ull;
constexpr uint32_t kHashTargetKeys = 192;   // at most 3/4 of the table's slots
constexpr uint32_t kBigTieSlots = 256;
constexpr uint32_t kStartMapWords = 8;
constexpr uint32_t kWaveTableWords = kBigTieSlots + kStartMapWords;
static_assert(kHashTargetKeys % 64u == 0u && kHashTargetKeys < kBigTieSlots,
              "whole chunks of 64 keys, and an empty slot for every probe sequence");
__host__ __device__ constexpr uint32_t target_lds(bool f64) {
    return ref_table_lds(f64) + (uint32_t)(kTargetBlock / 64) * kWaveTableWords * 8u;
}

// Count kernel (integer-valued columns in one pass).
constexpr int kCountBlockDefault = 256;     // threads per workgroup: 256, 512 or 1024
constexpr int kCountLdsDefault = 40448;     // dynamic LDS per workgroup: four per CU
constexpr int kCountLdsMax = 152 * 1024;    // largest dynamic LDS the tuning accepts
constexpr uint32_t kCountDummyWords = 64;      // LDS words the read pass's no-op atomics hit
constexpr uint32_t kMaxSlices = 64;         // target slices per column, beyond: sort path
constexpr uint32_t kCountMaxGroups = 65535; // group ids are packed in 16 bits
constexpr size_t kMatrixPad = 16;           // entries past nnz in a resident matrix (rows -1,
                                            // values 0)

// Per-column plan: written by the split kernel (and TA by the reference-table
// kernel), read by the later kernels of the same job.
struct ColPlan {
    uint64_t kmin, kmax;   // reference key range (sort path)
    uint64_t TA;           // reference-only tie sum over nonzero values
    uint32_t na;           // reference nonzero values
    uint32_t refneg;       // reference nonzero values below zero
    int32_t vmin;          // histogram path: bin = value - vmin
    uint32_t hist;         // 1 = histogram path, 0 = sort path
    uint32_t nb;           // histogram bins, or reference buckets (sort path)
    uint32_t shift;        // sort path: bucket = (key - kmin) >> shift
    uint32_t maxb;         // sort path, on-chip reference: keys in its fullest bucket
    uint32_t bigt;         // 4-byte keys: the column has targets of more than 64 keys,
                           // left by the window pass to the sorted pass
};

// A resident matrix's cached reference tables, per column (sort path, not the
// histogram branch): the sorted reference keys at rc_keys[koff ..
// koff + na), the bucket index at rc_tab[toff .. toff + nb], the reference's
// own tie sum and its fullest bucket.  They depend only on the column's values
// and the set of reference rows, so a later job on the same matrix with the
// same reference rows reads them instead of sorting the reference again.
struct RefCacheCol {
    int64_t koff;
    int64_t toff;
    uint64_t TA;
    uint64_t kmin, kmax;   // the reference's key range
    uint32_t na;
    uint32_t nb;
    uint32_t maxb;
    uint32_t shift;        // bucket = (key - kmin) >> shift
    uint32_t refneg;       // reference values below zero
    uint32_t valid;        // 1: the entries above hold this column's tables
};

// One Mann-Whitney job: a device-resident CSC matrix, its group labels and
// options, the per-job workspace of the kernel pipeline, and its outputs.
struct JobDesc {
    const void* vals;            // float* (f32 and small-int inputs) or double*
    const int32_t* rows;         // nnz row indices, -1 = out of range
    const int64_t* indptr;       // C+1
    const int32_t* gid;          // R group labels
    const uint32_t* gidp;        // R+1 group labels packed at gid_bits bits (pack_labels)
    uint32_t gid_bits;
    uint32_t gid_words;          // 32-bit words of gidp
    const uint32_t* group_rows;  // G rows per group
    // the rows in groups 0 .. G-1, ascending, and their groups: n_members =
    // the sum of group_rows.  Set when the matrix has columns storing rows
    // 0 .. R-1 in order, whose split gathers these rows' values instead of
    // reading every entry (or null)
    const uint32_t* members;
    const uint32_t* member_group;
    uint32_t n_members;
    const int32_t* colrange;     // 2C: value range of integer-valued columns (matrix metadata)
    const uint8_t* colident;     // C: 1 = the column stores rows 0 .. R-1 in order
    const uint8_t* colint;       // C: 1 = every stored value is a small integer (or non-
                                 // finite); 0 = the column cannot take a histogram branch
    const uint32_t* packed;      // nnz packed entries (pack_entry) or null; CountItem flags
    uint32_t pack_vb;            // value bits of a packed entry
    // workspace
    void* keys;                  // nnz keys (uint32 for float values, uint64 for double)
    uint32_t* cnt;               // C*G stored nonzero finite values per (column, group)
    uint32_t* zcnt;              // C*G stored explicit zeros per (column, group)
    uint32_t* goff;              // C*G start of each group's keys, from indptr[c]
    ColPlan* plan;               // C
    uint32_t* reftab;            // C*kTabStride reference tables
    void* spill;                 // keys (as `keys`) for reference slices beyond the on-chip
                                 // sort: the merge passes' second buffer
    const int64_t* spill_off;    // C: column c's first spill key, -1 = none (or null)
    // the matrix's cached reference tables for this job's reference rows (or
    // null): a valid column skips the reference sort and is ranked from them
    const RefCacheCol* rc;
    const void* rc_keys;
    const uint32_t* rc_tab;
    // outputs (gene-major, c*T + g-1); tie, n1o and n2o may be null (not kept)
    double* U1;
    double* P;
    double* tie;
    double* n1o;
    double* n2o;
    unsigned long long* err_key; // min over (column, rank) of recorded errors
    uint32_t* status;            // kStatus* bits
    unsigned long long* status_col;  // lowest column that raised a status bit
    int64_t C;
    int32_t R;
    int32_t G;                   // n_targets + 1
    int32_t T;
    int32_t mode;                // ZeroMode
    int32_t alt;                 // 0 less, 1 greater, 2 two-sided
    int32_t tie_corr;
    int32_t asym;                // 1: asymptotic p written with U1
    int32_t hist_max;            // largest integer range binned (0 = off)
    int32_t split_arrival;       // 1: the split keeps a sparse column's entry groups in LDS
                                 // and re-reads only its values to scatter
    uint32_t ref_cap;            // sort path: reference slices of at most this many keys
                                 // are sorted in LDS, larger ones in HBM (a power of two,
                                 // at most kRefKeysF32 / kRefKeysF64)
    uint32_t min_n;              // a group with fewer observations raises "Sample too small"
                                 // (2; 0 for the numba pipeline, which has no such check)
    double cc;                   // continuity correction, 0.5 or 0
};

struct WorkItem {
    int32_t job;
    int32_t col;
};

// One workgroup of the count kernel: column `col` (its entries [p0, p0+n),
// its histogram bins: values vmin .. vmin+nb-1) and targets t0+1 ..
// t0+min(S, T-t0).  S, the column's slice size, is the same in every slice of
// the column (T <= 65534, so both fit 16 bits).  The host fills the column
// fields from the matrix's own indptr and value-range metadata.
struct CountItem {
    int64_t p0;
    uint32_t n;
    int32_t vmin;   // smallest stored nonzero value (0 when there is none); for a packed
                    // item, its distinct values below zero (bins below the zero block)
    uint32_t nb;    // vmax - vmin + 1 (1 when there is no nonzero value); for a packed item,
                    // its distinct nonzero values (at least 1)
    int32_t col;
    uint16_t t0;
    uint16_t S;
    uint32_t flags;   // kCountPacked: read the column's packed entries; bits 8.. its job
                      // in a batched launch
};
constexpr uint32_t kCountPacked = 1u;
constexpr uint32_t kCountJobShift = 8u;
// words of the pack kernel's presence bitmap: every 16-bit integer value
constexpr uint32_t kPackBitmapWords = 2048u;

// Packed entries of a resident matrix.  A column whose finite nonzero values
// are all integers of one range [lo, lo + bins) (the columns the count kernel
// takes) keeps each entry also as one 32-bit word (row << vb) | code, with
// code = the value's rank among the column's distinct nonzero values (the
// count kernel's bin), the number of those values (at least 1: the count
// kernel's zero column) for an explicit zero and pack_skip_code for a
// non-finite value, and a row outside [0, R) stored as R.  The count
// kernel then reads one word per entry instead of a row and a value.  The row
// takes rb = pack_row_bits(R) bits (R itself included), the code vb = 32 - rb.
__host__ __device__ inline uint32_t pack_row_bits(uint32_t R) {
    uint32_t b = 1;
    while (b < 32u && (R >> b) != 0u) ++b;
    return b;
}
__host__ __device__ inline uint32_t pack_skip_code(uint32_t vb) { return (1u << vb) - 1u; }
// bins (0 or less: no nonzero value) and the zero code leave the skip code free
__host__ __device__ inline bool pack_fits(uint32_t vb, int64_t bins) {
    return vb >= 2u && vb <= 30u && bins <= (int64_t)(1u << vb) - 2;
}

// LDS of a count-kernel workgroup with `rows` rows (the reference and a
// slice's targets) over NB value bins plus a zero column:
//   dummy[kCountDummyWords] | acc[2*rows] u64 | cez, cnz, grs [rows] u32
//   (rounded to whole 8 B) | bins[NB] of (lt', eq) u32 pairs | counters[rows x rs]
// Counters are 16-bit (two per word) unless `wide` (a column of 2^16 entries
// or more: a count could reach 2^16).  The host's slicing and the kernel's
// own check use these same functions.
__host__ __device__ inline uint32_t count_row_stride(uint32_t nb1, bool wide) {
    // counters per row: an odd number of words, so threads that each walk
    // one row hit different banks
    if (wide) return nb1 | 1u;
    uint32_t r = (nb1 + 1u) & ~1u;
    if (((r >> 1) & 1u) == 0u) r += 2u;
    return r;
}
__host__ __device__ inline uint32_t count_row_words(uint32_t rows) { return (3u * rows + 1u) & ~1u; }
__host__ __device__ inline uint32_t count_cell_words(uint32_t rows, uint32_t rs, bool wide) {
    return wide ? rows * rs : (rows * rs + 1u) / 2u;
}
__host__ __device__ inline uint64_t count_lds_need(uint32_t rows, uint32_t NB, bool wide) {
    const uint32_t rs = count_row_stride(NB + 1u, wide);
    return 4ull * kCountDummyWords + 16ull * rows + 4ull * count_row_words(rows) + 8ull * NB +
           4ull * (uint64_t)count_cell_words(rows, rs, wide);
}

// Group labels for the count kernel's gathers: rows 0..R-1 and one more,
// row R (where rows out of range are sent), packed at `bits` bits each, least
// significant first within 32-bit words (a label may straddle two words; the
// kernel reads the 8 bytes from the label's first word).  A row with no group
// (label outside [0, G)) and row R hold 2^bits - 1, which is >= G.
inline uint32_t label_bits(uint32_t G) {
    uint32_t b = 1;
    while (((1u << b) - 1u) < G) ++b;
    return b;
}
inline size_t label_words(size_t R, uint32_t bits) { return ((R + 1) * bits + 31) / 32 + 2; }
inline void pack_labels(const int32_t* gid, size_t R, uint32_t G, uint32_t bits, uint32_t* out) {
    const size_t nw = label_words(R, bits);
    for (size_t w = 0; w < nw; ++w) out[w] = 0u;
    const uint32_t none = (1u << bits) - 1u;
    for (size_t r = 0; r <= R; ++r) {
        const uint32_t lab =
            (r < R && gid[r] >= 0 && (uint32_t)gid[r] < G) ? (uint32_t)gid[r] : none;
        const uint64_t pos = (uint64_t)r * bits;
        const size_t w = (size_t)(pos >> 5);
        const uint32_t sh = (uint32_t)(pos & 31u);
        out[w] |= lab << sh;
        if (sh + bits > 32u) out[w + 1] |= lab >> (32u - sh);
    }
}

// Slice size of a count-kernel column with T targets, NB bins and n stored
// entries in workgroups of lds_bytes: the largest S whose S + 1 rows fit,
// then evened out over the slices that takes.  0 when one target does not
// fit or the column would need more than kMaxSlices slices.
inline uint32_t count_slice_size(uint32_t T, uint32_t NB, int64_t n, uint32_t lds_bytes) {
    const bool wide = n >= 65536;
    if (T == 0u || T > 65534u || NB == 0u || NB > (uint32_t)kHistMaxBins ||
        count_lds_need(2u, NB, wide) > lds_bytes)
        return 0u;
    uint32_t lo = 1u, hi = T;   // need() grows with rows
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo + 1u) / 2u;
        if (count_lds_need(mid + 1u, NB, wide) <= lds_bytes) lo = mid;
        else hi = mid - 1u;
    }
    const uint32_t k = (T + lo - 1u) / lo;
    if (k > kMaxSlices) return 0u;
    return (T + k - 1u) / k;
}

// Exact-null table for one distinct (n1, n2): layout in the table buffer is
// [total][counts 0..U][cdf 0..U] doubles starting at `offset`.
struct ExactTable {
    int64_t offset;
    int32_t n1;
    int32_t n2;
};

struct ExactJob {
    const double* U1;
    double* P;
    const ExactTable* tables;
    const double* tabbuf;
    const int32_t* tidx_by_g;    // table per target (min/max/mix) or null
    const int32_t* tidx_by_pair; // table per pair (none) or null
    int64_t n_pairs;
    int32_t T;
    int32_t alt;
};

// The numba pipeline's p-values and target U (src/hpdex/backend/pykernel_.py),
// from the Mann-Whitney pass's reference U and raw tie sums (gene-major), into
// target-major outputs.
struct PdenJob {
    const double* U1;            // C*T reference U, c*T + t
    const double* tie;           // C*T sums of t^3 - t over tie runs (zero block included)
    const uint32_t* group_rows;  // G
    const uint8_t* exact;        // T*nchunk: 1 = the (target, gene chunk) takes the exact p
    const int32_t* tidx;         // T: exact table of the target, -1 = none
    const ExactTable* tables;
    const double* tabbuf;
    double* U_out;               // T*C target U, t*C + c
    double* P_out;               // T*C
    int64_t C;
    int64_t chunk;               // genes per decision chunk
    int64_t nchunk;
    int32_t T;
    int32_t tie_corr;
    double cc;                   // 0.5 or 0
};

// ---- launchers (hpdex_kernels.hip) ----
hipError_t launch_rows_i64_to_i32(const int64_t* in, int32_t* out, size_t n, int64_t R,
                                  hipStream_t s);
// convert any input dtype to the device value type (float or double)
hipError_t launch_convert_values(int dtype, const void* in, void* out, size_t n, hipStream_t s);
bool dtype_uses_double(int dtype);

// ---- ingest (hpdex_ingest.hip) ----
hipError_t launch_rows_i32_check(int32_t* rows, size_t n, int64_t R, hipStream_t s);
// Decode a compact upload in place (hpdex_ingest.hip): words[p] = (row & 0xFFFF)
// << 16 | value becomes the int32 row, vals[p] the float value; hi[c] holds
// the high bits of column c's first row.
hipError_t launch_decode_compact(uint32_t* words, float* vals, const int64_t* indptr,
                                 const int32_t* hi, size_t C, hipStream_t s);
size_t csr_workspace_bytes(size_t nnz);
// CSR (values of `dtype`, column indices and row pointers of index_bits 32 or
// 64, all on the device) -> CSC rows, values (float, or double when f64) and
// column pointers; *bad_host = 1 when a column index lies outside [0, C).
// Synchronises the stream.
hipError_t launch_csr_to_csc(int dtype, bool f64, const void* vals, const void* cols,
                             const void* indptr, int index_bits, size_t R, size_t C, size_t nnz,
                             void* vals_out, int32_t* rows_out, int64_t* indptr_out,
                             uint32_t* bad_host, void* ws, size_t ws_bytes, hipStream_t s);
// dense row-major R x C (values of `dtype` on the device) -> column pointers of
// its nonzeros (count phase, leaving a converted column-major copy at the
// start of ws), then rows and values (compact phase, reading that copy)
size_t dense_workspace_bytes(size_t R, size_t C, bool f64);
hipError_t launch_dense_count(int dtype, bool f64, const void* dense, size_t R, size_t C,
                              int64_t* indptr_out, void* ws, size_t ws_bytes, hipStream_t s);
hipError_t launch_dense_compact(bool f64, size_t R, size_t C, const int64_t* indptr, void* vals_out,
                                int32_t* rows_out, void* ws, hipStream_t s);

// One launch sequence of the Mann-Whitney pass of one job: the count kernel
// over the integer-valued columns the host routed to it (count_items, one
// workgroup per column and target slice), then the split -> reference ->
// target kernels over the other columns (gen_items[0 .. n_gen)).
struct PipelineLaunch {
    const JobDesc* jobs;       // device copy (split/reference/target kernels)
    JobDesc job;               // host copy, passed by value to the count kernel
    const CountItem* count_items; // workgroups of the count kernel (one job per launch)
    int n_count;                  // their number
    const WorkItem* gen_items;
    int n_gen;
    int G;                 // largest group count over the jobs (LDS sizing)
    bool f64;
    uint32_t count_lds;    // dynamic LDS bytes of the count kernel
    int count_block;       // its workgroup size, 256, 512 or 1024
    uint32_t split_lds_max;   // dynamic LDS the split kernel may take
};
// The batched launch of several jobs whose matrices share one value type:
// the count kernel over count_items (every job's, ordered largest first; each
// names its job), one workgroup per item, then the split -> reference ->
// target kernels over gen_items (every job's other columns, largest first,
// WorkItem.job naming the job).
struct BatchLaunch {
    const JobDesc* jobs;          // device array, indexed by an item's job
    const CountItem* count_items;
    int n_count;
    const WorkItem* gen_items;
    int n_gen;
    int G;                        // largest group count over the jobs
    bool f64;
    uint32_t count_lds;
    int count_block;
    uint32_t split_lds_max;
};
hipError_t launch_mwu_batch(const BatchLaunch& L, hipStream_t s);
// ev (optional, 5 events): recorded before the count kernel and after the
// count, split, reference and target kernels, for per-kernel timing.
hipError_t launch_mwu_pipeline(const PipelineLaunch& L, hipStream_t s,
                               hipEvent_t* ev = nullptr);
// Copy a finished job's reference tables of columns cols[0 .. n) into a
// matrix's cache (rc[c] holds their offsets, na, nb and TA; the kernel adds
// the fullest bucket and marks the entry valid).  `job` is the job's
// descriptor on the device; it must have run to completion on s.
hipError_t launch_refcache_fill(const JobDesc* job, bool f64, const int32_t* cols, uint32_t n,
                                RefCacheCol* rc, void* rc_keys, uint32_t* rc_tab, hipStream_t s);
// One stage of the sequence (0 count, 1 split, 2 reference, 3 target) alone.
hipError_t launch_mwu_stage(const PipelineLaunch& L, int stage, hipStream_t s);
// Allow the count kernel `bytes` of dynamic LDS and the split kernel
// kSplitLdsMax (once per process).
hipError_t prepare_count_kernel(uint32_t bytes);
// Column metadata of a resident matrix: ok[c] = 1 when every stored value of
// column c is non-finite or an integer in [-32768, 32767]; range[2c..2c+1] =
// smallest and largest stored nonzero finite value of such a column.
hipError_t launch_classify_columns(const void* vals, bool f64, const int64_t* indptr, int64_t C,
                                   uint8_t* ok, int32_t* range, const int32_t* rows, uint32_t R,
                                   uint8_t* ident, hipStream_t s);
hipError_t launch_pack_columns(const void* vals, bool f64, const int32_t* rows,
                               const int64_t* indptr, int64_t C, const int32_t* lo,
                               const int32_t* range, uint32_t R, uint32_t vb, uint32_t* out,
                               uint32_t* dense, uint32_t* zneg, hipStream_t s);
hipError_t launch_init_job_state(unsigned long long* err_key, uint32_t* status,
                                 unsigned long long* status_col, hipStream_t s);
hipError_t launch_exact_tables(const ExactTable* tabs, int n_tabs, double* tabbuf,
                               double* scratch, size_t scratch_per_table, hipStream_t s);
hipError_t launch_exact_p(const ExactJob& E, hipStream_t s);
hipError_t launch_group_mean(const void* vals, bool f64, const int32_t* rows,
                             const int64_t* indptr, const int32_t* gid,
                             const uint32_t* group_rows, int64_t C, int32_t R, int32_t G,
                             int include_zeros, double* out, hipStream_t s);
hipError_t launch_stream_copy(const void* in, void* out, size_t bytes, int grid, hipStream_t s);

// ---- DE post-processing (hpdex_de.hip) ----
// fc/log2fc[t*C + c] from group means[c*(T+1) + g] (de.py:126-142)
hipError_t launch_fold_change(const double* means, int64_t C, int32_t T, double clip, double* fc,
                              double* log2fc, hipStream_t s);
// Benjamini-Hochberg FDR of n p-values (n < 2^32) in a workspace of
// fdr_workspace_bytes(n); NaN everywhere if a p-value is outside [0, 1].
size_t fdr_workspace_bytes(int64_t n);
// An upper bound of fdr_workspace_bytes(n) computed without the device.
size_t fdr_workspace_bound(int64_t n);
hipError_t launch_fdr_bh(const double* p, int64_t n, double* fdr, void* ws, size_t ws_bytes,
                         hipStream_t s);

// ---- the numba pipeline, pden (hpdex_de.hip) ----
// sums[c*G + g] of every (column, group), added in row order in float32
// (acc_f32: a float32 matrix) or float64, as numpy's axis-0 mean adds them
hipError_t launch_pden_group_sums(const void* vals, bool f64, bool acc_f32, const int32_t* rows,
                                  const int64_t* indptr, const int32_t* gid, int64_t C, int32_t R,
                                  int32_t G, double* sums, hipStream_t s);
// fc/log2fc[t*C + c] from those sums over group_rows[g] rows, in float32
// arithmetic as numpy does for a float32 matrix (mean_f32) or in float64
// (src/hpdex/backend/de_.py:189-247, 640-652)
hipError_t launch_pden_fold_change(const double* sums, const uint32_t* group_rows, int64_t C,
                                   int32_t T, double clip, bool mean_f32, double* fc,
                                   double* log2fc, hipStream_t s);
// The per-(target, gene chunk) choice between the exact and the asymptotic p
// (pykernel_.py:258-262, 486-490; the histogram kernel's fallback of
// pykernel_.py:467-481 when use_hist), into exact[T*nchunk].  kmin/kmax:
// nchunk*G workspace (use_hist only); tflag: T*nchunk workspace.
hipError_t launch_pden_decide(const void* vals, bool f64, const int32_t* rows,
                              const int64_t* indptr, const int32_t* gid,
                              const uint32_t* group_rows, int64_t C, int32_t R, int32_t T,
                              const double* tie, int64_t chunk, int force, bool use_hist,
                              double max_bins, unsigned long long* kmin, unsigned long long* kmax,
                              uint32_t* tflag, uint8_t* exact, hipStream_t s);
hipError_t launch_pden_finish(const PdenJob& P, hipStream_t s);

}  // namespace hpdex
