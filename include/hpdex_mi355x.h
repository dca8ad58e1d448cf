/*
 * hpdex_mi355x.h -- C-ABI of the MI355X Mann-Whitney U differential-expression
 * backend.  Plain pointers and sizes only; every function returns 0 on success
 * or nonzero with a NUL-terminated message copied into err (err may be NULL).
 *
 * Host entry points (hpdex_mwu_csc, hpdex_group_mean_csc) take host buffers,
 * borrow them for the duration of the call, and write caller-allocated host
 * outputs.  They are reentrant: each call owns its stream and device memory.
 *
 * Device-resident entry points (hpdex_matrix_*, hpdex_plan_*) keep a CSC matrix
 * in HBM across calls for the batched form and for serving.
 */
#ifndef HPDEX_MI355X_H
#define HPDEX_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Element types of the CSC data array.  The reference dispatches the same ten
 * types (src/cpp/include/macro.hpp:4-14, src/cpp/module.cpp:144-166) and
 * promotes every value to double before comparing (src/cpp/mannwhitneyu.cpp:920). */
enum hpdex_dtype {
    HPDEX_INT8 = 0, HPDEX_UINT8 = 1, HPDEX_INT16 = 2, HPDEX_UINT16 = 3,
    HPDEX_INT32 = 4, HPDEX_UINT32 = 5, HPDEX_INT64 = 6, HPDEX_UINT64 = 7,
    HPDEX_FLOAT32 = 8, HPDEX_FLOAT64 = 9
};

/* Option block.  Enum values are those of MannWhitneyuOption
 * (src/cpp/include/mannwhitneyu.hpp:16-32):
 *   zero_handling: 0 none, 1 min, 2 max, 3 mix
 *   alternative:   0 less, 1 greater, 2 two-sided
 *   method:        1 exact, 2 asymptotic
 * ref_sorted / tar_sorted are accepted hints (results are defined only when
 * they hold; the device path sorts regardless).  fast_norm is accepted and the
 * precise erfc path is used for both values. */
typedef struct hpdex_mwu_opts {
    int32_t ref_sorted;
    int32_t tar_sorted;
    int32_t tie_correction;
    int32_t use_continuity;
    int32_t fast_norm;
    int32_t zero_handling;
    int32_t alternative;
    int32_t method;
} hpdex_mwu_opts;

/* Mann-Whitney U of every (gene column c, target group g) pair of a CSC matrix
 * against reference group 0.  Replaces hpdex::mannwhitneyu<T>
 * (src/cpp/include/mannwhitneyu.hpp:34-42; src/cpp/mannwhitneyu.cpp:1067-1111)
 * as bound by kernel.mannwhitneyu (src/cpp/module.cpp:35-107, 196-228).
 *   data/indices/indptr : CSC triple (nnz values of type dtype, nnz int64 row
 *                         indices, n_cols+1 int64 column pointers)
 *   group_id            : n_rows int32; 0 = reference, 1..n_targets = targets,
 *                         anything else = ignored row
 *   U1_out, P_out       : n_cols*n_targets doubles each, gene-major
 *                         (index c*n_targets + g-1)
 *   threads             : accepted for signature parity; the device path
 *                         ignores it.                                        */
int hpdex_mwu_csc(int dtype, const void* data, const int64_t* indices, const int64_t* indptr,
                  size_t n_rows, size_t n_cols, size_t nnz, const int32_t* group_id,
                  size_t n_targets, const hpdex_mwu_opts* opts, int threads,
                  double* U1_out, double* P_out, char* err, size_t err_len);

/* hpdex_mwu_csc without any device allocation: no counterpart in the
 * reference.  For callers that cannot call hipMalloc (tight serving loops,
 * harnesses that forbid allocation).  Same arguments, checks, messages and
 * results as hpdex_mwu_csc; every device buffer the call needs (the uploaded
 * matrix and the job's workspace) is carved from a fixed 64 MiB scratch
 * array that each GPU the library runs on holds until the process ends.  A
 * job whose bound, computed from its shape before anything runs (its values
 * and rows, about 34 KiB of tables per column, 16 bytes per pair), exceeds
 * the scratch is refused with an error naming both sizes; so is exact p in
 * mode none whose data-sized null tables do not fit.  Calls on one device share its
 * scratch and are serialised by a lock; the scratch keeps the last call's
 * contents, and no kernel reads a byte of it that the same call has not
 * written first.                                                            */
int hpdex_mwu_csc_noalloc(int dtype, const void* data, const int64_t* indices,
                          const int64_t* indptr, size_t n_rows, size_t n_cols, size_t nnz,
                          const int32_t* group_id, size_t n_targets, const hpdex_mwu_opts* opts,
                          int threads, double* U1_out, double* P_out, char* err,
                          size_t err_len);

/* hpdex_mwu_csc plus the per-pair statistics the reference keeps internally
 * (tie_arr, n1_arr, n2_arr of src/cpp/mannwhitneyu.cpp:859-861): tie_out is
 * the tie term sum(t^3 - t) (0 when tie_correction is off), n1_out/n2_out the
 * sample sizes.  Any of tie_out/n1_out/n2_out may be NULL.                   */
int hpdex_mwu_csc_stats(int dtype, const void* data, const int64_t* indices,
                        const int64_t* indptr, size_t n_rows, size_t n_cols, size_t nnz,
                        const int32_t* group_id, size_t n_targets, const hpdex_mwu_opts* opts,
                        int threads, double* U1_out, double* P_out, double* tie_out,
                        double* n1_out, double* n2_out, char* err, size_t err_len);

/* Per-(column, group) mean.  Replaces hpdex::group_mean<T>
 * (src/cpp/include/mannwhitneyu.hpp:45-53; src/cpp/mannwhitneyu.cpp:1113-1304)
 * as bound by kernel.group_mean (src/cpp/module.cpp:110-141, 230-242).
 * out: n_cols*n_groups doubles, index c*n_groups + g.                        */
int hpdex_group_mean_csc(int dtype, const void* data, const int64_t* indices,
                         const int64_t* indptr, size_t n_rows, size_t n_cols, size_t nnz,
                         const int32_t* group_id, size_t n_groups, int include_zeros,
                         int use_kahan, int threads, double* out, char* err, size_t err_len);

/* ---------------- differential-expression post-processing ----------------
 * Benjamini-Hochberg FDR of n p-values on the device; replaces
 * scipy.stats.false_discovery_control(P, method="bh") as called by
 * src/hpdex/de.py:172-184 (including its fallback: if a p-value is NaN or
 * outside [0, 1], every output is NaN). */
int hpdex_fdr_bh(const double* p, size_t n, double* fdr_out, char* err, size_t err_len);

/* The numeric part of parallel_differential_expression (src/hpdex/de.py:120-184)
 * in one device pass: group means with zeros (include_zeros=True) of the
 * reference and every target, fold change and log2 fold change with the clip
 * rules of de.py:136-141, the Mann-Whitney U test with zero_handling "min"
 * (two-sided, asymptotic), and the BH FDR of every p-value.  Layouts are
 * de.py's: fc_out / log2fc_out are n_targets x n_cols target-major (t*C + c);
 * U1_out, P_out, fdr_out are n_cols x n_targets gene-major (c*T + t), as
 * mannwhitneyu returns them.  Any output pointer may be NULL. */
int hpdex_de_csc(int dtype, const void* data, const int64_t* indices, const int64_t* indptr,
                 size_t n_rows, size_t n_cols, size_t nnz, const int32_t* group_id,
                 size_t n_targets, int tie_correction, int use_continuity, double clip_value,
                 int threads, double* U1_out, double* P_out, double* fc_out,
                 double* log2fc_out, double* fdr_out, char* err, size_t err_len);

/* Upper bound of the device memory one hpdex_de_csc call allocates for a
 * matrix of this shape (upload and staging, the test's workspace, group
 * means, fold changes, FDR), in bytes; 0 for an invalid dtype or n_targets 0.
 * Computed on the host: needs no device.  No counterpart in the reference;
 * hpdex_amd.stream sizes its gene runs under memory_limit_gb with it, as the
 * reference's streaming driver budgets its batches (src/hpdex/stream.py:161-262). */
size_t hpdex_de_csc_device_bytes(int dtype, size_t n_rows, size_t n_cols, size_t nnz,
                                 size_t n_targets);

/* ---------------- device-resident matrices and batched plans ----------------
 * No counterpart in the reference (its API is one call per job).  They let a
 * caller keep a CSC matrix in HBM and run several target sets against it.   */
typedef struct hpdex_matrix hpdex_matrix;
typedef struct hpdex_plan hpdex_plan;

/* Upload a host CSC matrix (same argument meaning as hpdex_mwu_csc). */
int hpdex_matrix_create(int dtype, const void* data, const int64_t* indices,
                        const int64_t* indptr, size_t n_rows, size_t n_cols, size_t nnz,
                        hpdex_matrix** out, char* err, size_t err_len);
/* A resident matrix built on the device from the layouts callers hold, with
 * 32- or 64-bit indices.  No counterpart in the reference: it replaces the
 * host conversions before the call -- scipy's tocsc of AnnData's CSR X and
 * csc_matrix of a dense X (src/hpdex/de.py:103-109), and the int64 copies of
 * indices and indptr (src/hpdex/backend/mannwhitneyu_.py:59-62).  The result
 * is the matrix hpdex_matrix_create would build from that CSC: rows in order
 * inside each column, explicit zeros of sparse input kept, zeros of dense
 * input (and -0.0) dropped, NaN kept.
 *   HPDEX_LAYOUT_CSC:   indices = nnz rows, indptr = n_cols+1 offsets
 *   HPDEX_LAYOUT_CSR:   indices = nnz columns, indptr = n_rows+1 offsets
 *                       (a column index outside [0, n_cols) is an error)
 *   HPDEX_LAYOUT_DENSE: data = n_rows*n_cols values, row-major; indices,
 *                       indptr and nnz unused
 * index_bits: 32 or 64, the width of indices and indptr.                    */
#define HPDEX_LAYOUT_CSC 0
#define HPDEX_LAYOUT_CSR 1
#define HPDEX_LAYOUT_DENSE 2
typedef struct hpdex_matrix_src {
    int32_t layout;
    int32_t dtype;
    const void* data;
    const void* indices;
    const void* indptr;
    int32_t index_bits;
    int32_t reserved;
    size_t n_rows, n_cols, nnz;
} hpdex_matrix_src;
int hpdex_matrix_create_from(const hpdex_matrix_src* src, hpdex_matrix** out, char* err,
                             size_t err_len);
/* Shape and contents of a resident matrix as CSC: values as double (nnz),
 * row indices (nnz; -1 marks rows outside [0, n_rows) of the input), column
 * pointers (n_cols+1).  Any output pointer may be NULL.                      */
int hpdex_matrix_shape(const hpdex_matrix* m, size_t* n_rows, size_t* n_cols, size_t* nnz);
int hpdex_matrix_download(const hpdex_matrix* m, double* data_out, int32_t* rows_out,
                          int64_t* indptr_out, char* err, size_t err_len);
/* hpdex_de_csc on a resident matrix, whatever layout it was built from (the
 * driver's steps 4-7 after its step 3 ran on the device).  No counterpart in
 * the reference; outputs and layouts as hpdex_de_csc.                        */
int hpdex_de_matrix(hpdex_matrix* m, const int32_t* group_id, size_t n_targets,
                    int tie_correction, int use_continuity, double clip_value, double* U1_out,
                    double* P_out, double* fc_out, double* log2fc_out, double* fdr_out, char* err,
                    size_t err_len);
/* The numba pipeline, hpdex.pden (src/hpdex/backend/de_.py:508-658 with
 * rank_sum_chunk_kernel_float / _hist of src/hpdex/backend/pykernel_.py:220-
 * 509), on a resident matrix built from the AnnData X: zeros rank at their
 * true position (the pipeline densifies X, de_.py:177-180), n is the group's
 * row count, and no group is too small for the test.  Every output is
 * target-major (t*n_cols + c, the row order of de_.py's DataFrame): U_out is
 * the TARGET group's U (pykernel_.py:74-76, 394-396), P_out its two-sided p
 * by pykernel_.py:84-147 or the exact 2 sf[rint(max U)] (:281-287), chosen
 * per (target, chunk of chunk_genes genes) as each task chooses it
 * (:267-271, 489-493; de_.py:113-124 and 291-333 make the tasks);
 * fc/log2fc by de_.py:189-247 on numpy's means (:640-652); fdr the BH of all
 * p-values (de_.py:250-288, 656).                                            */
typedef struct hpdex_pden_opts {
    int32_t tie_correction;
    int32_t continuity_correction;
    int32_t use_asymptotic;    /* -1: None (auto per chunk), 0: exact, 1: asymptotic */
    int32_t use_hist;          /* the histogram kernel runs (prefer_hist_if_int, an integer X,
                                  num_workers > 1): a chunk whose value range vmax - vmin + 1
                                  passes max_bins takes the asymptotic p (pykernel_.py:467-481) */
    int64_t max_bins;
    int64_t chunk_genes;       /* genes per task (de_.py:291-340), >= 1 */
    double clip_value;
} hpdex_pden_opts;
int hpdex_pden_matrix(hpdex_matrix* m, const int32_t* group_id, size_t n_targets,
                      const hpdex_pden_opts* o, double* U_out, double* P_out, double* fc_out,
                      double* log2fc_out, double* fdr_out, char* err, size_t err_len);
void hpdex_matrix_destroy(hpdex_matrix* m);
/* Bytes of HBM the resident matrix occupies (its reference tables below not
 * included). */
size_t hpdex_matrix_device_bytes(const hpdex_matrix* m);
/* How the matrix's entries crossed PCIe: 1 = compactly (a float32 CSC source:
 * one 32-bit word per entry, the row's low 16 bits and the value as a 16-bit
 * integer, decoded on the device; an entry qualifies when its value is a
 * whole number in [0, 65535] with the sign bit clear and its row lies in
 * [0, n_rows), less than 2^16 above the previous row of its column, and the
 * columns holding one that does not -- *patched_columns of them -- are then
 * written as they are), 0 = as given (8 bytes per entry after the int64 ->
 * int32 row conversion: other dtypes, copies under 4 MB, or more than an
 * eighth of the entries in columns that do not qualify).  The resident
 * matrix is the same either way.  No counterpart in the reference.          */
int hpdex_matrix_upload_form(const hpdex_matrix* m, int64_t* patched_columns);
/* Reference tables kept by a resident matrix.  No counterpart in the
 * reference, which sorts each gene's reference slice once per call
 * (src/cpp/mannwhitneyu.cpp:928-929).  After the first run of a plan (or a
 * session job) on the matrix whose reference rows -- the rows labelled 0 --
 * had no tables yet, the matrix keeps the sorted reference values and bucket
 * index of every column the sort path ranked (not the histogram branch), and
 * later jobs with the same reference rows skip the reference sort for those
 * columns: a new target set against the same reference.  One set of
 * reference rows is kept per matrix (the latest replaces the one before; a
 * job still reading the old one keeps it alive), in device memory of its own
 * (not a session's arena), until the matrix is destroyed or the tables are
 * cleared.  hpdex_tuning.ref_cache = 0 turns keeping and using them off.
 * Results are the same with and without them, bit for bit.
 * columns / bytes: columns with tables and their device bytes (0 if none). */
int hpdex_matrix_ref_cache_info(const hpdex_matrix* m, int64_t* columns, size_t* bytes);
void hpdex_matrix_ref_cache_clear(hpdex_matrix* m);

/* A job on a resident matrix: group labels and options fixed, outputs held on
 * the device.  form: HPDEX_PLAN_PLAIN (the kernel sequence launched eagerly
 * over the columns in order) or HPDEX_PLAN_BATCHED (columns queued largest
 * first, the whole launch sequence captured once in a hipGraph and replayed),
 * optionally or-ed with HPDEX_PLAN_STATS to keep the tie term and sample
 * sizes of every pair for hpdex_plan_results (otherwise only U1 and P are
 * written, which saves 24 bytes of HBM writes per pair).                    */
#define HPDEX_PLAN_PLAIN 0
#define HPDEX_PLAN_BATCHED 1
#define HPDEX_PLAN_STATS 2
int hpdex_plan_create(hpdex_matrix* m, const int32_t* group_id, size_t n_targets,
                      const hpdex_mwu_opts* opts, int form, hpdex_plan** out,
                      char* err, size_t err_len);
void hpdex_plan_destroy(hpdex_plan* p);
/* Run the plan `iters` times back to back on its stream and wait.  If ms_total
 * is non-NULL it receives the elapsed milliseconds (HIP events on the plan's
 * stream); if kernel_ms is non-NULL it receives the mean duration of one run
 * of the device work (HIP events around each run).                          */
int hpdex_plan_run(hpdex_plan* p, int iters, double* ms_total, double* kernel_ms,
                   char* err, size_t err_len);
/* Copy the last run's results to host buffers (n_cols*n_targets doubles each;
 * any pointer may be NULL; tie_out, n1_out and n2_out need a plan created with
 * HPDEX_PLAN_STATS). Returns the run's error, if it had one.                 */
int hpdex_plan_results(hpdex_plan* p, double* U1_out, double* P_out, double* tie_out,
                       double* n1_out, double* n2_out, char* err, size_t err_len);

/* Where the plan's columns go: the one-pass count kernel (integer-valued
 * columns) or the split/reference/target kernels; handed_over is always 0
 * (a count-kernel column holding a value outside its recorded range makes
 * the run fail with an error naming the column).  count_blocks is the count
 * kernel's workgroups: one per column, or one per slice of targets when a
 * column's counters for every target do not fit one workgroup's LDS;
 * count_packed_columns the count-kernel columns read as packed entries.     */
typedef struct hpdex_plan_info {
    int64_t count_columns;
    int64_t count_nnz;
    int64_t split_columns;
    int64_t handed_over;
    int64_t count_blocks;
    int64_t count_packed_columns;
    int64_t ref_cached_columns;   /* sort-path columns whose reference tables come
                                     from the matrix's kept tables (no sort) */
} hpdex_plan_info;
int hpdex_plan_get_info(hpdex_plan* p, hpdex_plan_info* out, char* err, size_t err_len);
/* Per-kernel timing: runs the plain launch sequence `iters` times with HIP
 * events between its kernels (on the plan's stream) and writes the mean
 * milliseconds of the count, split, reference and target kernels to ms[0..3]
 * (a kernel with no columns times as an empty launch).  The results are
 * those of a normal run.                                                    */
int hpdex_plan_stage_times(hpdex_plan* p, int iters, double* ms, char* err, size_t err_len);
/* Mean duration of one kernel of the plan (0 count, 1 split, 2 reference, 3
 * target: its window and sorted passes) over `iters` launches of it back to
 * back on the plan's stream, between two HIP events, with one run of the
 * whole launch sequence before and one after (so the plan's results and
 * state are those of a normal run).  No counterpart in the reference.       */
int hpdex_plan_kernel_ms(hpdex_plan* p, int kernel, int iters, double* ms, char* err,
                         size_t err_len);

/* ---------------- batches: several plans in one launch sequence ----------------
 * No counterpart in the reference.  A batch runs several plans (jobs on
 * resident matrices: different target sets on one matrix, or jobs on
 * different matrices) as one launch sequence captured in a hipGraph: the
 * count kernel over one queue of every plan's (column, target slice) items,
 * ordered largest first and handed out to the CUs in that order, then the
 * split/reference/target kernels over every plan's other columns (largest
 * first), then the exact p lookups.  Each plan's results stay in the plan
 * (hpdex_plan_results).  The plans are borrowed for the batch's life and
 * must not run on their own while a batch run is in flight.  Plans of exact
 * p in zero_handling none cannot join.                                      */
typedef struct hpdex_batch hpdex_batch;
int hpdex_batch_create(hpdex_plan* const* plans, size_t n_plans, hpdex_batch** out, char* err,
                       size_t err_len);
/* Replay the batch `iters` times back to back and wait; ms_total receives
 * the elapsed milliseconds (HIP events on the batch's stream). */
int hpdex_batch_run(hpdex_batch* b, int iters, double* ms_total, char* err, size_t err_len);
/* The batch's count-kernel items (workgroups) and sort-path columns. */
int hpdex_batch_get_info(hpdex_batch* b, int64_t* count_items, int64_t* split_columns);
void hpdex_batch_destroy(hpdex_batch* b);

/* Tunable choices of the device path.  hpdex_amd.autotune times candidate
 * settings on a caller's job and applies the fastest through
 * hpdex_set_tuning; the defaults are its choice on integer-count matrices.  */
typedef struct hpdex_tuning {
    int32_t hist_max_bins;     /* widest integer value range a column may have to use
                                  the histogram path (0 = never, at most 8192)        */
    int32_t size_ordered;      /* batched form: 1 = queue the largest columns first  */
    int32_t count_lds_bytes;   /* LDS per workgroup of the one-pass kernel for integer-
                                  valued columns (0 = off: every column takes the
                                  split/reference/target kernels; else 16 KiB..152 KiB) */
    int32_t count_block;       /* threads per workgroup of that kernel: 256, 512, 1024 */
    int32_t count_packed;      /* 1 = that kernel reads a matrix's packed entries (one
                                  32-bit word per entry, made when the matrix is
                                  created) where it has them; 0 = rows and values  */
    int32_t split_arrival;     /* sort path: 1 = the split kernel keeps each entry's
                                  group and key from its counting pass in LDS
                                  (columns whose entries fit) and scatters the keys
                                  from there; 0 = counting and scattering passes
                                  both read rows, values and group labels           */
    int32_t ref_lds_keys;      /* sort path, the LDS/global sort cutover: a column's
                                  reference slice of at most this many values is
                                  sorted in LDS, a larger one in HBM (a power of two
                                  in [64, 8192]; 8-byte values: at most 4096 on chip) */
    int32_t ref_cache;         /* 1 = resident matrices keep the sorted reference of
                                  sort-path jobs for later jobs with the same
                                  reference rows (hpdex_matrix_ref_cache_info); 0 =
                                  every job sorts its reference                    */
} hpdex_tuning;
int hpdex_get_tuning(hpdex_tuning* out);
int hpdex_set_tuning(const hpdex_tuning* t, char* err, size_t err_len);

/* ---------------- serving session ----------------
 * No counterpart in the reference (its API is one blocking call per job).
 * A session serves DE jobs of different sizes concurrently on one GPU: it owns
 * one HBM arena of hbm_budget_bytes and `slots` worker threads, each with its
 * own stream.  Submitted jobs wait in FIFO order and are admitted as soon as a
 * slot is free and the job's footprint (an upper bound of everything it
 * allocates: its uploaded matrix, if any, and its workspace) fits in the
 * arena; admitted jobs upload and set up concurrently, and the kernels of
 * every job set up at the time run in one merged launch (hpdex_session_batches).  A job either names a resident matrix (hpdex_matrix_create; a new
 * target set against cells already in HBM) or carries a host CSC that is
 * uploaded into its arena block and dropped when it ends.  All pointers in a
 * job are borrowed until hpdex_session_wait returns for it.                  */
typedef struct hpdex_session hpdex_session;
typedef struct hpdex_session_opts {
    size_t hbm_budget_bytes;   /* arena size */
    int32_t slots;             /* jobs in flight at once (1..64) */
} hpdex_session_opts;
typedef struct hpdex_job {
    hpdex_matrix* matrix;      /* resident matrix, or NULL to upload the host matrix below */
    int32_t dtype;             /* host matrix (same meaning as hpdex_mwu_csc) */
    const void* data;
    const void* indices;       /* int64, or int32 with index_bits 32 */
    const void* indptr;
    size_t n_rows, n_cols, nnz;
    const int32_t* group_id;   /* n_rows labels (of the matrix's rows) */
    size_t n_targets;
    hpdex_mwu_opts opts;       /* method exact needs zero_handling min, max or mix */
    double* U1_out;            /* n_cols * n_targets each, written when the job ends */
    double* P_out;
    int32_t index_bits;        /* 0 or 64: int64 indices and indptr; 32: int32 */
    int32_t layout;            /* HPDEX_LAYOUT_CSC (0) or HPDEX_LAYOUT_CSR, converted on the
                                  device as hpdex_matrix_create_from does */
} hpdex_job;
typedef struct hpdex_job_times {
    double queue_ms;           /* submit -> admitted */
    double run_ms;             /* admitted -> results in the host buffers */
    double latency_ms;         /* submit -> results in the host buffers */
    size_t footprint_bytes;    /* arena bytes the job held */
    /* run_ms in phases: */
    double upload_ms;          /* the matrix the job brings, to the device and classified (0
                                  for a resident matrix) */
    double setup_ms;           /* labels, routing, work lists and exact tables */
    double kernels_ms;         /* set up -> the merged launch holding its kernels ended */
    double finish_ms;          /* exact p lookup, errors, results to the host buffers */
} hpdex_job_times;
int hpdex_session_create(const hpdex_session_opts* o, hpdex_session** out, char* err,
                         size_t err_len);
/* Queue a job; *ticket identifies it.  Fails at once (nothing queued) when the
 * job can never fit the arena or its arguments are invalid. */
int hpdex_session_submit(hpdex_session* s, const hpdex_job* job, int64_t* ticket, char* err,
                         size_t err_len);
/* Block until the job has ended; returns its error (0 on success) and, if
 * times is not NULL, its timings.  Each ticket is waited for once. */
int hpdex_session_wait(hpdex_session* s, int64_t ticket, hpdex_job_times* times, char* err,
                       size_t err_len);
/* Largest number of jobs that ran at once, and most arena bytes in use, so far. */
int hpdex_session_peak(hpdex_session* s, int32_t* jobs_in_flight, size_t* arena_bytes);
/* Continuous batching so far: merged kernel launches, the jobs they ran, and
 * the most jobs one launch held.  Every job's kernels run in the launch the
 * session's launcher makes of all jobs set up (uploaded, labels and work
 * lists on the device) at the time; jobs set up while it runs form the next. */
int hpdex_session_batches(hpdex_session* s, int64_t* launches, int64_t* jobs, int64_t* max_jobs);
/* Wait for every queued job, then free the arena and the workers. */
void hpdex_session_destroy(hpdex_session* s);

/* Attainable HBM bandwidth: a 16-byte streaming copy of `bytes` (read + write
 * counted), run `iters` times; gbs receives GB/s.  Also the calibration kernel
 * for HBM byte counters (its traffic is exactly 2*bytes per launch). */
int hpdex_bandwidth_probe(size_t bytes, int iters, double* gbs, char* err, size_t err_len);

/* Library version string and device summary, for logs. */
const char* hpdex_version(void);
int hpdex_device_info(int* n_cus, int* max_lds_per_block, char* name, size_t name_len);

#ifdef __cplusplus
}
#endif

#endif /* HPDEX_MI355X_H */
