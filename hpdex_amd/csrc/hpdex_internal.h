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
    kStatusCapacity = 1u,   // a slice exceeded the workgroup's LDS budget
    kStatusGroups = 2u,     // too many groups for the LDS counters
};

// One Mann-Whitney job: a device-resident CSC matrix, its group labels and
// options, and where its per-pair results go.  Lives in device memory for the
// persistent kernel (one entry per job in flight) or is passed by value.
struct JobDesc {
    const void* vals;            // float* (f32 and small-int inputs) or double*
    const int32_t* rows;         // nnz row indices, -1 = out of range
    const int64_t* indptr;       // C+1
    const int32_t* gid;          // R group labels
    const uint32_t* group_rows;  // G rows per group
    double* U1;                  // C*T
    double* P;                   // C*T (written when fuse_p)
    double* tie;                 // C*T or null
    double* n1o;                 // C*T or null
    double* n2o;                 // C*T or null
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
    int32_t fuse_p;              // asymptotic p computed in the epilogue
    int32_t hist_max;            // largest integer range binned (0 = off)
    double cc;                   // continuity correction, 0.5 or 0
};

struct WorkItem {
    int32_t job;
    int32_t col;
};

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

// ---- launchers (hpdex_kernels.hip) ----
hipError_t launch_rows_i64_to_i32(const int64_t* in, int32_t* out, size_t n, int64_t R,
                                  hipStream_t s);
// convert any input dtype to the device value type (float or double)
hipError_t launch_convert_values(int dtype, const void* in, void* out, size_t n, hipStream_t s);
bool dtype_uses_double(int dtype);

hipError_t launch_gene_plain(const JobDesc& J, bool f64, int block, int lds, hipStream_t s);
hipError_t launch_gene_persistent(const JobDesc* jobs, const WorkItem* items, int n_items,
                                  unsigned* head, unsigned* done, bool f64, int grid, int block,
                                  int lds, hipStream_t s);
hipError_t launch_init_job_state(unsigned long long* err_key, uint32_t* status,
                                 unsigned long long* status_col, hipStream_t s);
hipError_t launch_exact_tables(const ExactTable* tabs, int n_tabs, double* tabbuf,
                               double* scratch, size_t scratch_per_table, hipStream_t s);
hipError_t launch_exact_p(const ExactJob& E, hipStream_t s);
hipError_t launch_group_mean(const void* vals, bool f64, const int32_t* rows,
                             const int64_t* indptr, const int32_t* gid,
                             const uint32_t* group_rows, int64_t C, int32_t R, int32_t G,
                             int include_zeros, double* out, uint32_t* status, int block,
                             int lds, hipStream_t s);
hipError_t set_gene_kernel_lds_limit(int lds);

}  // namespace hpdex
