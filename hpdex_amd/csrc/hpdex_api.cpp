// hpdex_api.cpp -- host side of the MI355X Mann-Whitney U backend: the C-ABI
// declared in include/hpdex_mi355x.h, device memory, streams, the size-ordered
// work queue, exact-null table planning and hipGraph capture.
//
// Host-side checks mirror the reference's:
//   n_targets >= 1                 src/cpp/mannwhitneyu.cpp:1088-1090
//   unsupported dtype              src/cpp/module.cpp:165-166
//   per-column sample-size errors  src/cpp/mannwhitneyu.cpp:938-1038, raised
//   after the pass (:1045); the device records the lowest (column, group).
#include "hpdex_mi355x.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <set>
#include <string>
#include <vector>

#include "hpdex_host.h"
#include "hpdex_internal.h"

using namespace hpdex;
using namespace hpdex_host;

namespace hpdex_host {

std::mutex g_mu;
Tuning g_tune;
bool g_init = false;
int g_cus = 0;
int g_max_lds = 0;
bool g_count_ready = false;   // the count kernel may use up to kCountLdsMax of LDS
char g_name[256] = {0};
thread_local BumpBlock* t_block = nullptr;

int set_err(char* err, size_t len, const std::string& m) {
    if (err && len) {
        std::strncpy(err, m.c_str(), len - 1);
        err[len - 1] = 0;
    }
    return 1;
}

int ensure_init(char* err, size_t err_len) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_init) return 0;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return set_err(err, err_len,
                       "no HIP device available: the MI355X backend has no CPU fallback");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t pr;
    HIPCHK(hipGetDeviceProperties(&pr, dev));
    g_cus = pr.multiProcessorCount;
    g_max_lds = (int)pr.sharedMemPerBlock;
    std::snprintf(g_name, sizeof g_name, "%s (%s)", pr.name, pr.gcnArchName);
    g_count_ready = g_max_lds >= kCountLdsMax && prepare_count_kernel(kCountLdsMax) == hipSuccess;
    if (!g_count_ready) g_tune.count_lds = 0;
    g_init = true;
    return 0;
}

Tuning tuning() {
    std::lock_guard<std::mutex> lk(g_mu);
    return g_tune;
}

bool valid_dtype(int d) { return d >= HPDEX_INT8 && d <= HPDEX_FLOAT64; }
size_t dtype_size(int d) {
    switch (d) {
    case HPDEX_INT8: case HPDEX_UINT8: return 1;
    case HPDEX_INT16: case HPDEX_UINT16: return 2;
    case HPDEX_INT32: case HPDEX_UINT32: case HPDEX_FLOAT32: return 4;
    default: return 8;
    }
}

// Column pointers must be a valid CSC index before any kernel reads through
// them (the reference trusts them; a bad one would fault the device here).
int check_indptr(const int64_t* indptr, size_t C, size_t nnz, char* err, size_t err_len) {
    if (indptr[0] < 0) return set_err(err, err_len, "indptr[0] must be >= 0");
    for (size_t c = 0; c < C; ++c)
        if (indptr[c + 1] < indptr[c])
            return set_err(err, err_len, "indptr must be non-decreasing");
    if ((size_t)indptr[C] > nnz)
        return set_err(err, err_len, "indptr[-1] exceeds the number of stored values");
    return 0;
}

// ---------------------------------------------------------------------------
// Resident matrix upload, per-job state
// ---------------------------------------------------------------------------
// A float32 CSC source in 4 bytes per entry (upload_compact) when its values
// are small whole numbers and its columns' rows ascend with small gaps,
// decoded in place into the matrix's rows and values; the few columns that do
// not qualify are then written as they are.  `done` says whether the compact
// form was taken.  m->R, C, nnz, rows, vals and indptr are set up.
template <typename IDX>
static int upload_compact_csc(hpdex_matrix* m, const float* vals, const IDX* rows,
                              const int64_t* indptr, bool& done, hipStream_t s, char* err,
                              size_t err_len) {
    done = false;
    DevBuf hi;
    HIPCHK(hi.alloc(m->C * 4));
    std::vector<int32_t> patch;
    const hipError_t e = upload_compact<IDX>(m->rows.as<uint32_t>(), hi.as<int32_t>(), vals, rows,
                                             indptr, m->C, m->nnz, (int64_t)m->R, s, patch);
    if (e == hipErrorNotSupported) return 0;
    HIPCHK(e);
    HIPCHK(hipMemcpyAsync(m->indptr.p, indptr, (m->C + 1) * 8, hipMemcpyHostToDevice, s));
    HIPCHK(launch_decode_compact(m->rows.as<uint32_t>(), (float*)m->vals.p,
                                 m->indptr.as<int64_t>(), hi.as<int32_t>(), m->C, s));
    // the columns that did not qualify, as they are (rows outside [0, R) -> -1)
    std::vector<int32_t> prow;
    size_t total = 0;
    for (int32_t c : patch) total += (size_t)(indptr[c + 1] - indptr[c]);
    prow.reserve(total);
    for (int32_t c : patch)
        for (int64_t p = indptr[c]; p < indptr[c + 1]; ++p) {
            const int64_t r = (int64_t)rows[p];
            prow.push_back((r >= 0 && r < (int64_t)m->R) ? (int32_t)r : -1);
        }
    size_t at = 0;
    for (int32_t c : patch) {
        const int64_t p0 = indptr[c], n = indptr[c + 1] - p0;
        if (n <= 0) continue;
        HIPCHK(hipMemcpyAsync(m->rows.as<int32_t>() + p0, prow.data() + at, (size_t)n * 4,
                              hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync((float*)m->vals.p + p0, vals + p0, (size_t)n * 4,
                              hipMemcpyHostToDevice, s));
        at += (size_t)n;
    }
    HIPCHK(hipStreamSynchronize(s));   // hi and prow are freed on return
    done = true;
    m->compact = true;
    m->compact_patched = (int64_t)patch.size();
    return 0;
}

int upload_matrix(hpdex_matrix* m, int dtype, const void* data, const int64_t* indices,
                  const int64_t* indptr, size_t R, size_t C, size_t nnz, hipStream_t s,
                  char* err, size_t err_len) {
    if (!valid_dtype(dtype)) return set_err(err, err_len, "Unsupported dtype for data");
    if (R > (size_t)INT32_MAX) return set_err(err, err_len, "more than 2^31-1 rows is unsupported");
    if (C > (size_t)INT32_MAX) return set_err(err, err_len, "more than 2^31-1 columns is unsupported");
    if (check_indptr(indptr, C, nnz, err, err_len)) return 1;
    m->dtype = dtype;
    m->f64 = dtype_uses_double(dtype);
    m->R = R;
    m->C = C;
    m->nnz = nnz;
    const size_t vsz = m->f64 ? 8 : 4;
    // kMatrixPad extra entries (rows -1, values 0) past the last column
    HIPCHK(m->vals.alloc((nnz + kMatrixPad) * vsz));
    HIPCHK(m->rows.alloc((nnz + kMatrixPad) * 4));
    HIPCHK(hipMemsetAsync((char*)m->vals.p + nnz * vsz, 0, kMatrixPad * vsz, s));
    HIPCHK(hipMemsetAsync((char*)m->rows.p + nnz * 4, 0xFF, kMatrixPad * 4, s));
    HIPCHK(m->indptr.alloc((C + 1) * 8));
    bool compact = false;
    if (nnz && dtype == HPDEX_FLOAT32)
        if (upload_compact_csc<int64_t>(m, (const float*)data, indices, indptr, compact, s, err,
                                        err_len))
            return 1;
    if (nnz && !compact) {
        if (dtype == HPDEX_FLOAT32 || dtype == HPDEX_FLOAT64) {
            HIPCHK(upload_bytes(m->vals.p, data, nnz * vsz, s));
        } else {
            DevBuf raw;
            HIPCHK(raw.alloc(nnz * dtype_size(dtype)));
            HIPCHK(upload_bytes(raw.p, data, nnz * dtype_size(dtype), s));
            HIPCHK(launch_convert_values(dtype, raw.p, m->vals.p, nnz, s));
            HIPCHK(hipStreamSynchronize(s));
        }
        // the int64 rows become int32 on the host while they are staged (half
        // the bytes over PCIe), or on the device after a plain copy
        const hipError_t e = upload_rows_i64(m->rows.as<int32_t>(), indices, nnz, (int64_t)R, s);
        if (e == hipErrorNotSupported) {
            DevBuf idx64;
            HIPCHK(idx64.alloc(nnz * 8));
            HIPCHK(hipMemcpyAsync(idx64.p, indices, nnz * 8, hipMemcpyHostToDevice, s));
            HIPCHK(launch_rows_i64_to_i32(idx64.as<int64_t>(), m->rows.as<int32_t>(), nnz,
                                          (int64_t)R, s));
        } else {
            HIPCHK(e);
        }
        HIPCHK(hipStreamSynchronize(s));
    }
    HIPCHK(hipMemcpyAsync(m->indptr.p, indptr, (C + 1) * 8, hipMemcpyHostToDevice, s));
    m->h_indptr.assign(indptr, indptr + C + 1);
    return finish_matrix(m, s, err, err_len);
}

// Column metadata of a matrix whose values, rows and column pointers are on
// the device (h_indptr set): integer-valued columns and their value ranges,
// which let the count kernel bin values during its one read of a column, and
// the columns by size.
int finish_matrix(hpdex_matrix* m, hipStream_t s, char* err, size_t err_len) {
    const size_t C = m->C;
    const int64_t* indptr = m->h_indptr.data();
    m->col_int.assign(C, 0);
    m->h_range.assign(2 * C, 0);
    HIPCHK(m->colrange.alloc(2 * C * 4));
    HIPCHK(m->colident.alloc(C ? C : 1));
    HIPCHK(m->colint.alloc(C ? C : 1));
    if (C) {
        HIPCHK(launch_classify_columns(m->vals.p, m->f64, m->indptr.as<int64_t>(), (int64_t)C,
                                       m->colint.as<uint8_t>(), m->colrange.as<int32_t>(),
                                       m->rows.as<int32_t>(), (uint32_t)m->R,
                                       m->colident.as<uint8_t>(), s));
        HIPCHK(hipMemcpyAsync(m->col_int.data(), m->colint.p, C, hipMemcpyDeviceToHost, s));
        std::vector<uint8_t> ident(C);
        HIPCHK(hipMemcpyAsync(ident.data(), m->colident.p, C, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        m->any_ident = std::find(ident.begin(), ident.end(), (uint8_t)1) != ident.end();
        HIPCHK(hipMemcpyAsync(m->h_range.data(), m->colrange.p, 2 * C * 4, hipMemcpyDeviceToHost,
                              s));
        HIPCHK(hipStreamSynchronize(s));
    }
    // packed entries of the integer-valued columns whose bins fit the code
    // bits: one 32-bit word per entry for the count kernel's read
    m->col_packed.assign(C, 0);
    m->pack_vb = 32u - pack_row_bits((uint32_t)m->R);
    std::vector<int32_t> lo(C, INT_MIN);
    bool any = false;
    for (size_t c = 0; c < C; ++c) {
        const int64_t bins = (int64_t)m->h_range[2 * c + 1] - m->h_range[2 * c] + 1;
        if (m->col_int[c] && pack_fits(m->pack_vb, bins)) {
            lo[c] = bins > 0 ? m->h_range[2 * c] : 0;
            m->col_packed[c] = 1;
            any = true;
        }
    }
    if (any && m->nnz && m->pack) {
        DevBuf d_lo, d_dense, d_zneg;
        HIPCHK(d_lo.alloc(C * 4));
        HIPCHK(d_dense.alloc(C * 4));
        HIPCHK(d_zneg.alloc(C * 4));
        HIPCHK(hipMemcpyAsync(d_lo.p, lo.data(), C * 4, hipMemcpyHostToDevice, s));
        HIPCHK(m->packed.alloc(m->nnz * 4));
        HIPCHK(launch_pack_columns(m->vals.p, m->f64, m->rows.as<int32_t>(),
                                   m->indptr.as<int64_t>(), (int64_t)C, d_lo.as<int32_t>(),
                                   m->colrange.as<int32_t>(), (uint32_t)m->R, m->pack_vb,
                                   m->packed.as<uint32_t>(), d_dense.as<uint32_t>(),
                                   d_zneg.as<uint32_t>(), s));
        m->h_dense.assign(C, 0);
        m->h_zneg.assign(C, 0);
        HIPCHK(hipMemcpyAsync(m->h_dense.data(), d_dense.p, C * 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(m->h_zneg.data(), d_zneg.p, C * 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));   // the staging buffers are freed on return
    } else {
        m->col_packed.assign(C, 0);
    }
    m->order.resize(C);
    std::iota(m->order.begin(), m->order.end(), 0);
    std::stable_sort(m->order.begin(), m->order.end(), [&](int32_t a, int32_t b) {
        return (indptr[a + 1] - indptr[a]) > (indptr[b + 1] - indptr[b]);
    });
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

// Row or column pointers of a source layout, widened to int64 on the host and
// checked as a CSC index is.
int widen_indptr(const void* p, int bits, size_t n, size_t nnz, std::vector<int64_t>& out,
                        char* err, size_t err_len) {
    out.resize(n + 1);
    for (size_t i = 0; i <= n; ++i)
        out[i] = bits == 32 ? (int64_t)((const int32_t*)p)[i] : ((const int64_t*)p)[i];
    return check_indptr(out.data(), n, nnz, err, err_len);
}

// Values of the source's dtype on the device, converted to the matrix's value
// type later by the kernels of each layout.
static hipError_t upload_raw(DevBuf& b, const void* host, size_t bytes, hipStream_t s) {
    hipError_t e = b.alloc(bytes);
    if (e != hipSuccess || bytes == 0) return e;
    return upload_bytes(b.p, host, bytes, s);
}

int upload_matrix_from(hpdex_matrix* m, const hpdex_matrix_src* src, hipStream_t s, char* err,
                       size_t err_len) {
    const int dtype = src->dtype;
    const size_t R = src->n_rows, C = src->n_cols;
    if (!valid_dtype(dtype)) return set_err(err, err_len, "Unsupported dtype for data");
    if (R > (size_t)INT32_MAX) return set_err(err, err_len, "more than 2^31-1 rows is unsupported");
    if (C > (size_t)INT32_MAX) return set_err(err, err_len, "more than 2^31-1 columns is unsupported");
    const int bits = src->index_bits;
    if (src->layout != HPDEX_LAYOUT_DENSE && bits != 32 && bits != 64)
        return set_err(err, err_len, "index_bits must be 32 or 64");
    if (src->layout == HPDEX_LAYOUT_CSC && bits == 64)
        return upload_matrix(m, dtype, src->data, (const int64_t*)src->indices,
                             (const int64_t*)src->indptr, R, C, src->nnz, s, err, err_len);
    m->dtype = dtype;
    m->f64 = dtype_uses_double(dtype);
    m->R = R;
    m->C = C;
    const size_t vsz = m->f64 ? 8 : 4;
    auto alloc_out = [&](size_t nnz) -> hipError_t {
        m->nnz = nnz;
        hipError_t e;
        if ((e = m->vals.alloc((nnz + kMatrixPad) * vsz)) != hipSuccess) return e;
        if ((e = m->rows.alloc((nnz + kMatrixPad) * 4)) != hipSuccess) return e;
        if ((e = hipMemsetAsync((char*)m->vals.p + nnz * vsz, 0, kMatrixPad * vsz, s)) != hipSuccess)
            return e;
        if ((e = hipMemsetAsync((char*)m->rows.p + nnz * 4, 0xFF, kMatrixPad * 4, s)) != hipSuccess)
            return e;
        return m->indptr.alloc((C + 1) * 8);
    };
    if (src->layout == HPDEX_LAYOUT_CSC) {   // 32-bit indices
        const size_t nnz = src->nnz;
        if (widen_indptr(src->indptr, 32, C, nnz, m->h_indptr, err, err_len)) return 1;
        HIPCHK(alloc_out(nnz));
        bool compact = false;
        if (nnz && dtype == HPDEX_FLOAT32)
            if (upload_compact_csc<int32_t>(m, (const float*)src->data,
                                            (const int32_t*)src->indices, m->h_indptr.data(),
                                            compact, s, err, err_len))
                return 1;
        if (nnz && !compact) {
            if (dtype == HPDEX_FLOAT32 || dtype == HPDEX_FLOAT64) {
                HIPCHK(upload_bytes(m->vals.p, src->data, nnz * vsz, s));
            } else {
                DevBuf raw;
                HIPCHK(upload_raw(raw, src->data, nnz * dtype_size(dtype), s));
                HIPCHK(launch_convert_values(dtype, raw.p, m->vals.p, nnz, s));
                HIPCHK(hipStreamSynchronize(s));
            }
            HIPCHK(upload_bytes(m->rows.p, src->indices, nnz * 4, s));
            HIPCHK(launch_rows_i32_check(m->rows.as<int32_t>(), nnz, (int64_t)R, s));
        }
        HIPCHK(hipMemcpyAsync(m->indptr.p, m->h_indptr.data(), (C + 1) * 8, hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
        return finish_matrix(m, s, err, err_len);
    }
    if (src->layout == HPDEX_LAYOUT_CSR) {
        const size_t nnz = src->nnz;
        if (nnz > (size_t)UINT32_MAX) return set_err(err, err_len, "more than 2^32-1 stored values in CSR input");
        std::vector<int64_t> rptr;
        if (widen_indptr(src->indptr, bits, R, nnz, rptr, err, err_len)) return 1;
        if ((size_t)rptr[R] != nnz || rptr[0] != 0)
            return set_err(err, err_len, "CSR indptr must start at 0 and end at nnz");
        HIPCHK(alloc_out(nnz));
        DevBuf raw, cols, ptr, ws;
        HIPCHK(upload_raw(raw, src->data, nnz * dtype_size(dtype), s));
        HIPCHK(upload_raw(cols, src->indices, nnz * (size_t)(bits / 8), s));
        HIPCHK(upload_raw(ptr, src->indptr, (R + 1) * (size_t)(bits / 8), s));
        const size_t wsb = csr_workspace_bytes(nnz);
        HIPCHK(ws.alloc(wsb));
        uint32_t bad = 0;
        HIPCHK(launch_csr_to_csc(dtype, m->f64, raw.p, cols.p, ptr.p, bits, R, C, nnz, m->vals.p,
                                 m->rows.as<int32_t>(), m->indptr.as<int64_t>(), &bad, ws.p, wsb, s));
        if (bad) return set_err(err, err_len, "column index out of range in CSR input");
        m->h_indptr.resize(C + 1);
        HIPCHK(hipMemcpyAsync(m->h_indptr.data(), m->indptr.p, (C + 1) * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        return finish_matrix(m, s, err, err_len);
    }
    if (src->layout == HPDEX_LAYOUT_DENSE) {
        if (R && C > (size_t)-1 / R / 8) return set_err(err, err_len, "dense input too large");
        DevBuf raw, ws, ptr;
        HIPCHK(upload_raw(raw, src->data, R * C * dtype_size(dtype), s));
        const size_t wsb = dense_workspace_bytes(R, C, m->f64);
        HIPCHK(ws.alloc(wsb));
        HIPCHK(ptr.alloc((C + 1) * 8));
        HIPCHK(launch_dense_count(dtype, m->f64, raw.p, R, C, ptr.as<int64_t>(), ws.p, wsb, s));
        m->h_indptr.resize(C + 1);
        HIPCHK(hipMemcpyAsync(m->h_indptr.data(), ptr.p, (C + 1) * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        raw.reset();
        const size_t nnz = (size_t)m->h_indptr[C];
        HIPCHK(alloc_out(nnz));
        HIPCHK(hipMemcpyAsync(m->indptr.p, ptr.p, (C + 1) * 8, hipMemcpyDeviceToDevice, s));
        HIPCHK(launch_dense_compact(m->f64, R, C, ptr.as<int64_t>(), m->vals.p, m->rows.as<int32_t>(),
                                    ws.p, s));
        HIPCHK(hipStreamSynchronize(s));
        return finish_matrix(m, s, err, err_len);
    }
    return set_err(err, err_len, "unknown layout");
}

std::vector<uint64_t> reference_rows(const int32_t* group_id, size_t R) {
    std::vector<uint64_t> bits((R + 63) / 64, 0ull);
    for (size_t r = 0; r < R; ++r)
        if (group_id[r] == 0) bits[r >> 6] |= 1ull << (r & 63);
    return bits;
}

namespace {
// DevBuf allocations in this scope take device memory of their own
// (hipMalloc) even on a session worker, whose job buffers come from its block.
struct OwnAllocations {
    BumpBlock* saved = t_block;
    OwnAllocations() { t_block = nullptr; }
    ~OwnAllocations() { t_block = saved; }
};
}  // namespace

int refcache_fill(JobState& js, const hpdex_matrix& m, hipStream_t s, char* err, size_t err_len) {
    if (!js.rc_fill) return 0;
    js.rc_fill = false;
    unsigned long long h[3] = {0, 0, 0};
    HIPCHK(hipMemcpyAsync(h, js.flags.p, 24, hipMemcpyDeviceToHost, s));
    std::vector<ColPlan> plans(js.C);
    if (js.C) HIPCHK(hipMemcpyAsync(plans.data(), js.plan.p, js.C * sizeof(ColPlan),
                                    hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (h[0] != ~0ull || (uint32_t)h[2] != 0u) return 0;   // a failed job leaves no tables
    auto e = std::make_shared<RefCacheEntry>();
    e->ref_rows = js.ref_rows;
    e->valid.assign(js.C, 0);
    std::vector<RefCacheCol> cols(js.C);
    std::memset(cols.data(), 0, cols.size() * sizeof(RefCacheCol));
    std::vector<int32_t> list;
    int64_t koff = 0, toff = 0;
    for (const WorkItem& wi : js.h_gen_items) {
        const ColPlan& pl = plans[wi.col];
        if (pl.hist) continue;   // histogram tables depend on the targets' values too
        RefCacheCol& rc = cols[wi.col];
        rc.koff = koff;
        rc.toff = toff;
        rc.TA = pl.TA;
        rc.kmin = pl.kmin;
        rc.kmax = pl.kmax;
        rc.shift = pl.shift;
        rc.refneg = pl.refneg;
        rc.na = pl.na;
        rc.nb = pl.nb;
        koff += pl.na;
        toff += (int64_t)pl.nb + 1;
        list.push_back(wi.col);
        e->valid[wi.col] = 1;
    }
    if (list.empty()) return 0;
    const size_t ksz = m.f64 ? 8 : 4;
    DevBuf d_list;
    {
        OwnAllocations own;
        HIPCHK(e->cols.alloc(js.C * sizeof(RefCacheCol)));
        HIPCHK(e->keys.alloc((size_t)koff * ksz));
        HIPCHK(e->tab.alloc((size_t)toff * 4));
        HIPCHK(d_list.alloc(list.size() * 4));
    }
    HIPCHK(hipMemcpyAsync(e->cols.p, cols.data(), js.C * sizeof(RefCacheCol),
                          hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d_list.p, list.data(), list.size() * 4, hipMemcpyHostToDevice, s));
    HIPCHK(launch_refcache_fill(js.d_job.as<JobDesc>(), m.f64, d_list.as<int32_t>(),
                                (uint32_t)list.size(), e->cols.as<RefCacheCol>(), e->keys.p,
                                e->tab.as<uint32_t>(), s));
    HIPCHK(hipStreamSynchronize(s));
    e->n_cols = (int64_t)list.size();
    e->bytes = e->cols.n + e->keys.n + e->tab.n;
    std::lock_guard<std::mutex> lk(m.rc_mu);
    m.rc = std::move(e);
    return 0;
}

int job_setup(JobState& js, const hpdex_matrix* m, const int32_t* group_id, size_t T,
              const hpdex_mwu_opts* o, bool stats, hipStream_t s, char* err, size_t err_len) {
    if (T < 1) return set_err(err, err_len, "[mannwhitney] n_targets must be >= 1 (>=1 target groups)");
    if (T + 1 > (size_t)INT32_MAX / 4) return set_err(err, err_len, "too many target groups");
    js.T = T;
    js.G = T + 1;
    js.C = m->C;
    js.opts = *o;
    js.stats = stats;
    js.group_rows.assign(js.G, 0u);
    js.rc_use.reset();
    js.rc_fill = false;
    js.rc_cols = 0;
    js.ref_rows.clear();
    if (m->resident) js.ref_rows = reference_rows(group_id, m->R);
    for (size_t r = 0; r < m->R; ++r) {
        const int g = group_id[r];
        if (g >= 0 && (size_t)g < js.G) ++js.group_rows[g];
    }
    const size_t np = js.C * T;
    HIPCHK(js.gid.alloc(m->R * 4));
    if (m->R) HIPCHK(hipMemcpyAsync(js.gid.p, group_id, m->R * 4, hipMemcpyHostToDevice, s));
    // packed labels for the count kernel's gathers (it runs only when
    // G <= 65535 and every label's bit position fits 32 bits)
    js.gid_bits = 0;
    js.gid_words = 0;
    if (js.G <= kCountMaxGroups) {
        const uint32_t bits = label_bits((uint32_t)js.G);
        if ((uint64_t)(m->R + 1) * bits < ((uint64_t)1 << 32)) {
            js.gid_bits = bits;
            js.gid_words = label_words(m->R, bits);
            std::vector<uint32_t> packed(js.gid_words);
            pack_labels(group_id, m->R, (uint32_t)js.G, bits, packed.data());
            HIPCHK(js.gidp.alloc(js.gid_words * 4));
            HIPCHK(hipMemcpyAsync(js.gidp.p, packed.data(), js.gid_words * 4, hipMemcpyHostToDevice, s));
            HIPCHK(hipStreamSynchronize(s));   // packed is a host temporary
        }
    }
    HIPCHK(js.grows.alloc(js.G * 4));
    HIPCHK(hipMemcpyAsync(js.grows.p, js.group_rows.data(), js.G * 4, hipMemcpyHostToDevice, s));
    // the rows in a group, ascending, with their groups: the split of a
    // column in row order gathers their values (a wave's members span a few
    // hundred rows, so its gathers fall in a few cache lines)
    js.n_members = 0;
    if (m->any_ident) {
        uint32_t M = 0;
        for (size_t g = 0; g < js.G; ++g) M += js.group_rows[g];
        std::vector<uint32_t> mem(M ? M : 1), mgr(M ? M : 1);
        uint32_t at = 0;
        for (size_t r = 0; r < m->R; ++r) {
            const int g = group_id[r];
            if (g >= 0 && (size_t)g < js.G) {
                mem[at] = (uint32_t)r;
                mgr[at] = (uint32_t)g;
                ++at;
            }
        }
        HIPCHK(js.members.alloc((size_t)M * 4));
        HIPCHK(js.member_group.alloc((size_t)M * 4));
        if (M) {
            HIPCHK(hipMemcpyAsync(js.members.p, mem.data(), (size_t)M * 4, hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(js.member_group.p, mgr.data(), (size_t)M * 4,
                                  hipMemcpyHostToDevice, s));
        }
        HIPCHK(hipStreamSynchronize(s));   // mem and mgr are host temporaries
        js.n_members = M;
    }
    HIPCHK(js.U1.alloc(np * 8));
    HIPCHK(js.P.alloc(np * 8));
    // sample sizes are also needed by the exact p of zero_handling none
    const bool sizes = stats || (o->method == 1 && o->zero_handling == 0);
    if (stats) HIPCHK(js.tie.alloc(np * 8));
    if (sizes) {
        HIPCHK(js.n1.alloc(np * 8));
        HIPCHK(js.n2.alloc(np * 8));
    }
    HIPCHK(js.flags.alloc(32));
    // the sort path's workspace (keys, counts, offsets, column plans,
    // reference tables) is allocated by job_upload_desc, for jobs that route
    // a column there
    HIPCHK(launch_init_job_state(js.err_key(), js.status(), js.status_col(), s));
    return 0;
}

JobDesc make_desc(const JobState& js, const hpdex_matrix* m, const Tuning& tu) {
    JobDesc J{};
    J.vals = m->vals.p;
    J.rows = m->rows.as<int32_t>();
    J.indptr = m->indptr.as<int64_t>();
    J.gid = js.gid.as<int32_t>();
    J.gidp = js.gidp.as<uint32_t>();
    J.gid_bits = js.gid_bits;
    J.gid_words = (uint32_t)js.gid_words;
    J.group_rows = js.grows.as<uint32_t>();
    J.members = js.n_members ? js.members.as<uint32_t>() : nullptr;
    J.member_group = js.n_members ? js.member_group.as<uint32_t>() : nullptr;
    J.n_members = js.n_members;
    J.colrange = m->colrange.as<int32_t>();
    J.colident = m->colident.as<uint8_t>();
    J.colint = m->colint.as<uint8_t>();
    J.packed = m->packed.p ? m->packed.as<uint32_t>() : nullptr;
    J.pack_vb = m->pack_vb;
    J.keys = js.keys.p;
    J.cnt = js.cnt.as<uint32_t>();
    J.zcnt = js.zcnt.as<uint32_t>();
    J.goff = js.goff.as<uint32_t>();
    J.plan = js.plan.as<ColPlan>();
    J.reftab = js.reftab.as<uint32_t>();
    J.U1 = js.U1.as<double>();
    J.P = js.P.as<double>();
    J.tie = js.tie.as<double>();   // null when not kept
    J.n1o = js.n1.as<double>();
    J.n2o = js.n2.as<double>();
    J.err_key = js.err_key();
    J.status = js.status();
    J.status_col = js.status_col();
    J.C = (int64_t)m->C;
    J.R = (int32_t)m->R;
    J.G = (int32_t)js.G;
    J.T = (int32_t)js.T;
    J.mode = js.opts.zero_handling;
    J.alt = js.opts.alternative;
    J.tie_corr = js.opts.tie_correction ? 1 : 0;
    J.asym = (js.opts.method == 2) ? 1 : 0;
    J.hist_max = tu.hist_max;
    J.split_arrival = tu.split_arrival;
    J.ref_cap = (uint32_t)ref_lds_cap(m->f64, tu);
    J.min_n = 2u;
    J.cc = js.opts.use_continuity ? 0.5 : 0.0;
    return J;
}

// Put the job descriptor and its work queues on the device: integer-valued
// columns to the count kernel, one workgroup per slice of targets whose
// counter rows fit its LDS (count_slice_size), the rest to the
// split/reference/target kernels; each queue in column order (plain form) or
// largest first (batched form).  Columns of 2^28 entries or more take the
// split path: the count kernel addresses a column with 32-bit byte offsets.
int job_upload_desc(JobState& js, const JobDesc& J, const hpdex_matrix& m, const Tuning& tu,
                    bool size_ordered, hipStream_t s, char* err, size_t err_len) {
    js.J = J;
    const bool count_ok = tu.count_lds > 0 && js.G >= 2 && js.G <= kCountMaxGroups && js.gid_bits > 0;
    const uint32_t T = (uint32_t)js.T;
    std::vector<CountItem> count_items;
    std::vector<WorkItem> gen_items;
    count_items.reserve(m.C);
    gen_items.reserve(m.C);
    js.n_count_cols = 0;
    js.n_count_packed = 0;
    js.count_nnz = 0;
    for (size_t k = 0; k < m.C; ++k) {
        const int32_t c = size_ordered ? m.order[k] : (int32_t)k;
        const int64_t len = m.h_indptr[c + 1] - m.h_indptr[c];
        // a packed column's bins are its distinct values; any other's its range
        const bool packed = tu.count_packed && m.col_packed[c];
        const int64_t bins = packed ? (int64_t)m.h_dense[c]
                                    : (int64_t)m.h_range[2 * c + 1] - m.h_range[2 * c] + 1;
        const uint32_t S =
            (count_ok && m.col_int[c] && bins <= tu.hist_max && len < ((int64_t)1 << 28))
                ? count_slice_size(T, (uint32_t)(bins > 0 ? bins : 1), len, (uint32_t)tu.count_lds)
                : 0u;
        if (S == 0u) {
            gen_items.push_back(WorkItem{0, c});
            continue;
        }
        CountItem it{};
        it.p0 = m.h_indptr[c];
        it.n = (uint32_t)len;
        it.vmin = packed ? (int32_t)m.h_zneg[c] : (bins > 0 ? m.h_range[2 * c] : 0);
        it.nb = bins > 0 ? (uint32_t)bins : 1u;
        it.col = c;
        it.S = (uint16_t)S;
        it.flags = packed ? kCountPacked : 0u;
        for (uint32_t t0 = 0; t0 < T; t0 += S) {
            it.t0 = (uint16_t)t0;
            count_items.push_back(it);
        }
        js.n_count_cols += 1;
        js.n_count_packed += it.flags ? 1 : 0;
        js.count_nnz += len;
    }
    js.n_count = (int)count_items.size();
    js.n_gen_static = (int)gen_items.size();
    if (!gen_items.empty()) {   // the sort path's workspace
        HIPCHK(js.keys.alloc(m.nnz * (m.f64 ? 8 : 4)));
        HIPCHK(js.cnt.alloc(js.C * js.G * 4));
        HIPCHK(js.zcnt.alloc(js.C * js.G * 4));
        HIPCHK(js.goff.alloc(js.C * js.G * 4));
        HIPCHK(js.plan.alloc(js.C * sizeof(ColPlan)));
        HIPCHK(js.reftab.alloc(js.C * (size_t)kTabStride * 4));
        js.J.keys = js.keys.p;
        js.J.cnt = js.cnt.as<uint32_t>();
        js.J.zcnt = js.zcnt.as<uint32_t>();
        js.J.goff = js.goff.as<uint32_t>();
        js.J.plan = js.plan.as<ColPlan>();
        js.J.reftab = js.reftab.as<uint32_t>();
        // the matrix's reference tables for these reference rows, if it has
        // them; otherwise this job fills them after its first run
        if (tu.ref_cache && m.resident) {
            std::shared_ptr<RefCacheEntry> e;
            {
                std::lock_guard<std::mutex> lk(m.rc_mu);
                e = m.rc;
            }
            if (e && e->ref_rows == js.ref_rows) {
                js.J.rc = e->cols.as<RefCacheCol>();
                js.J.rc_keys = e->keys.p;
                js.J.rc_tab = e->tab.as<uint32_t>();
                for (const WorkItem& wi : gen_items) js.rc_cols += e->valid[wi.col];
                js.rc_use = std::move(e);
            } else {
                js.rc_fill = true;
            }
        }
    }
    // spill space for the sort-path columns whose reference slice may pass
    // the on-chip sort (its stored values: at most the column's entries and
    // the reference's rows)
    {
        const size_t cap = js.J.ref_cap;
        std::vector<int64_t> off;
        int64_t total = 0;
        for (const WorkItem& wi : gen_items) {
            const size_t b = spill_keys_of(m.h_indptr[wi.col + 1] - m.h_indptr[wi.col],
                                           js.group_rows[0], cap);
            if (b == 0) continue;
            if (off.empty()) off.assign(m.C, -1);
            off[wi.col] = total;
            total += (int64_t)b;
        }
        js.J.spill = nullptr;
        js.J.spill_off = nullptr;
        if (!off.empty()) {
            HIPCHK(js.spill.alloc((size_t)total * (m.f64 ? 8 : 4)));
            HIPCHK(js.spill_off.alloc(m.C * 8));
            HIPCHK(hipMemcpyAsync(js.spill_off.p, off.data(), m.C * 8, hipMemcpyHostToDevice, s));
            HIPCHK(hipStreamSynchronize(s));   // off is a host temporary
            js.J.spill = js.spill.p;
            js.J.spill_off = js.spill_off.as<int64_t>();
        }
    }
    HIPCHK(js.d_job.alloc(sizeof(JobDesc)));
    HIPCHK(hipMemcpyAsync(js.d_job.p, &js.J, sizeof(JobDesc), hipMemcpyHostToDevice, s));
    js.count_lds = (uint32_t)tu.count_lds;
    js.count_block = tu.count_block;
    HIPCHK(js.d_count_items.alloc(count_items.size() * sizeof(CountItem)));
    if (!count_items.empty())
        HIPCHK(hipMemcpyAsync(js.d_count_items.p, count_items.data(),
                              count_items.size() * sizeof(CountItem), hipMemcpyHostToDevice, s));
    HIPCHK(js.d_gen_items.alloc(gen_items.size() * sizeof(WorkItem)));
    if (!gen_items.empty())
        HIPCHK(hipMemcpyAsync(js.d_gen_items.p, gen_items.data(),
                              gen_items.size() * sizeof(WorkItem), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    js.h_count_items = std::move(count_items);
    js.h_gen_items = std::move(gen_items);   // the item vectors are host temporaries
    return 0;
}

PipelineLaunch pipeline_of(const JobState& js, bool f64) {
    PipelineLaunch L{};
    L.jobs = js.d_job.as<JobDesc>();
    L.job = js.J;
    L.count_items = js.d_count_items.as<CountItem>();
    L.n_count = js.n_count;
    L.gen_items = js.d_gen_items.as<WorkItem>();
    L.n_gen = js.n_gen_static;
    L.G = (int)js.G;
    L.f64 = f64;
    L.count_lds = js.count_lds;
    L.count_block = js.count_block;
    L.split_lds_max = g_count_ready ? (uint32_t)kSplitLdsMax : (uint32_t)kLdsBudget;
    return L;
}

// Plan the exact-null tables.  min/max/mix: sizes are the group row counts,
// known now.  none: sizes depend on the data; the caller passes n1/n2 per
// pair read back after the U pass.
int exact_plan(JobState& js, const double* n1pair, const double* n2pair, hipStream_t s,
               char* err, size_t err_len) {
    std::map<std::pair<int32_t, int32_t>, int32_t> idx;
    js.tabs.clear();
    int64_t off = 0;
    size_t max_cells = 0;
    auto add = [&](int32_t a, int32_t b) -> int32_t {
        auto key = std::make_pair(a, b);
        auto it = idx.find(key);
        if (it != idx.end()) return it->second;
        const int64_t U = (int64_t)a * b;
        ExactTable t;
        t.offset = off;
        t.n1 = a;
        t.n2 = b;
        off += 1 + 2 * (U + 1);
        const size_t s_ = (size_t)std::min(a, b);
        const size_t cells = (s_ + 1) * (size_t)(U + 1);
        max_cells = std::max(max_cells, cells);
        const int32_t id = (int32_t)js.tabs.size();
        js.tabs.push_back(t);
        idx[key] = id;
        return id;
    };
    std::vector<int32_t> tidx;
    if (!n1pair) {
        js.exact_by_pair = false;
        js.tidx_by_g.resize(js.T);
        for (size_t g = 1; g <= js.T; ++g)
            js.tidx_by_g[g - 1] = add((int32_t)js.group_rows[0], (int32_t)js.group_rows[g]);
        tidx = js.tidx_by_g;
    } else {
        js.exact_by_pair = true;
        const size_t np = js.C * js.T;
        tidx.resize(np);
        for (size_t i = 0; i < np; ++i) tidx[i] = add((int32_t)n1pair[i], (int32_t)n2pair[i]);
    }
    if (max_cells > kExactMaxCells) {
        for (const auto& t : js.tabs) {
            const size_t s_ = (size_t)std::min(t.n1, t.n2);
            if ((s_ + 1) * ((size_t)t.n1 * t.n2 + 1) > kExactMaxCells)
                return set_err(err, err_len,
                               "exact null table too large for n1=" + std::to_string(t.n1) +
                                   ", n2=" + std::to_string(t.n2) + "; use method='asymptotic'");
        }
    }
    js.scratch_per_table = 2 * max_cells;
    HIPCHK(js.d_tabs.alloc(js.tabs.size() * sizeof(ExactTable)));
    HIPCHK(hipMemcpyAsync(js.d_tabs.p, js.tabs.data(), js.tabs.size() * sizeof(ExactTable),
                          hipMemcpyHostToDevice, s));
    HIPCHK(js.tabbuf.alloc((size_t)off * 8));
    HIPCHK(js.tscratch.alloc(js.tabs.size() * js.scratch_per_table * 8));
    HIPCHK(js.d_tidx.alloc(tidx.size() * 4));
    HIPCHK(hipMemcpyAsync(js.d_tidx.p, tidx.data(), tidx.size() * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

int exact_launch(const JobState& js, hipStream_t s, bool build_tables, char* err, size_t err_len) {
    if (build_tables)
        HIPCHK(launch_exact_tables(js.d_tabs.as<ExactTable>(), (int)js.tabs.size(),
                                   js.tabbuf.as<double>(), js.tscratch.as<double>(),
                                   js.scratch_per_table, s));
    ExactJob E{};
    E.U1 = js.U1.as<double>();
    E.P = js.P.as<double>();
    E.tables = js.d_tabs.as<ExactTable>();
    E.tabbuf = js.tabbuf.as<double>();
    E.tidx_by_g = js.exact_by_pair ? nullptr : js.d_tidx.as<int32_t>();
    E.tidx_by_pair = js.exact_by_pair ? js.d_tidx.as<int32_t>() : nullptr;
    E.n_pairs = (int64_t)(js.C * js.T);
    E.T = (int32_t)js.T;
    E.alt = js.opts.alternative;
    HIPCHK(launch_exact_p(E, s));
    return 0;
}

// Read the job's error state and turn it into the reference's message.
int job_errors(const JobState& js, hipStream_t s, char* err, size_t err_len) {
    unsigned long long h[3] = {0, 0, 0};
    HIPCHK(hipMemcpyAsync(h, js.flags.p, 24, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const unsigned long long key = h[0];
    const uint32_t status = (uint32_t)h[2];
    if (key != ~0ull) {
        const unsigned long long span = 2ull * js.G + 1ull;
        const unsigned long long col = key / span, rank = key % span;
        if (rank < js.G)
            return set_err(err, err_len,
                           "Non-canonical CSC: column " + std::to_string(col) +
                               " holds more entries for group " + std::to_string(rank) +
                               " than the group has rows");
        if (rank == js.G)
            return set_err(err, err_len,
                           "Sample too small for reference at column " + std::to_string(col));
        return set_err(err, err_len,
                       "Sample too small for group " + std::to_string(rank - js.G) +
                           " at column " + std::to_string(col));
    }
    if (status) {
        const std::string why =
            (status & kStatusGroups)
                ? "more target groups than the split workgroup's LDS holds counters for"
            : (status & kStatusRange)
                ? "a stored value outside the column's recorded value range (the resident "
                  "matrix's metadata does not match its values)"
                : "the reference slice passes the on-chip sort (8,192 4-byte or 4,096 8-byte "
                  "keys) and the job reserved no spill space for it";
        return set_err(err, err_len, "column " + std::to_string(h[1]) + ": " + why);
    }
    return 0;
}

int copy_out(const JobState& js, double* U1, double* P, double* tie, double* n1, double* n2,
             hipStream_t s, char* err, size_t err_len) {
    const size_t b = js.C * js.T * 8;
    if ((tie && !js.tie.p) || (n1 && !js.n1.p) || (n2 && !js.n2.p))
        return set_err(err, err_len,
                       "the tie term and sample sizes are kept only by plans created with "
                       "HPDEX_PLAN_STATS");
    if (U1 && b) HIPCHK(download_bytes(U1, js.U1.p, b, s));
    if (P && b) HIPCHK(download_bytes(P, js.P.p, b, s));
    if (tie && b) HIPCHK(download_bytes(tie, js.tie.p, b, s));
    if (n1 && b) HIPCHK(download_bytes(n1, js.n1.p, b, s));
    if (n2 && b) HIPCHK(download_bytes(n2, js.n2.p, b, s));
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

bool opts_ok(const hpdex_mwu_opts* o) {
    return o && o->zero_handling >= 0 && o->zero_handling <= 3 && o->alternative >= 0 &&
           o->alternative <= 2 && (o->method == 1 || o->method == 2);
}

// Exact p for zero_handling none needs the data-dependent sample sizes.
int run_exact_none(JobState& js, hipStream_t s, char* err, size_t err_len) {
    const size_t np = js.C * js.T;
    std::vector<double> n1(np), n2(np);
    if (np) {
        HIPCHK(hipMemcpyAsync(n1.data(), js.n1.p, np * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(n2.data(), js.n2.p, np * 8, hipMemcpyDeviceToHost, s));
    }
    HIPCHK(hipStreamSynchronize(s));
    if (exact_plan(js, n1.data(), n2.data(), s, err, err_len)) return 1;
    return exact_launch(js, s, true, err, err_len);
}

size_t up256(size_t n) { return ((n ? n : 16) + 255) & ~(size_t)255; }

size_t matrix_footprint(int dtype, size_t R, size_t C, size_t nnz, int layout, int index_bits,
                        bool pack) {
    const bool f64 = dtype_uses_double(dtype);
    const size_t vsz = f64 ? 8 : 4, ib = index_bits == 32 ? 4 : 8;
    // values and rows with their padding, column pointers, value ranges, the
    // in-order flags and the classification's staging flags
    size_t f = up256((nnz + kMatrixPad) * vsz) + up256((nnz + kMatrixPad) * 4) + up256((C + 1) * 8);
    f += up256(2 * C * 4) + 2 * up256(C);
    if (layout == HPDEX_LAYOUT_CSC && dtype == HPDEX_FLOAT32) f += up256(C * 4);   // compact upload
    if (layout == HPDEX_LAYOUT_CSR) {
        // raw values, column indices and row pointers, the transpose workspace
        f += up256(nnz * dtype_size(dtype)) + up256(nnz * ib) + up256((R + 1) * ib);
        f += up256(csr_workspace_bytes(nnz));
    } else {
        if (dtype != HPDEX_FLOAT32 && dtype != HPDEX_FLOAT64) f += up256(nnz * dtype_size(dtype));
        if (ib == 8) f += up256(nnz * 8);   // int64 rows staged for the conversion
    }
    if (pack) f += up256(nnz * 4) + 3 * up256(C * 4);
    return f;
}

size_t spill_keys_of(int64_t len, uint32_t gr0, size_t cap) {
    const size_t b = std::min((size_t)len, (size_t)gr0);
    return b > cap ? b : 0;
}

size_t ref_lds_cap(bool f64, const Tuning& tu) {
    const size_t hw = f64 ? (size_t)kRefKeysF64 : (size_t)kRefKeysF32;
    return std::min(hw, (size_t)tu.ref_lds_keys);
}

size_t spill_keys_bound(const int64_t* indptr, size_t C, size_t nnz, uint32_t gr0, bool f64,
                        const Tuning& tu) {
    const size_t cap = ref_lds_cap(f64, tu);
    if (gr0 <= cap) return 0;
    if (!indptr) return std::min(nnz, C * (size_t)gr0);
    size_t t = 0;
    for (size_t c = 0; c < C; ++c) t += spill_keys_of(indptr[c + 1] - indptr[c], gr0, cap);
    return t;
}

size_t job_footprint(size_t R, size_t C, size_t nnz, bool f64, size_t T, const hpdex_mwu_opts& o,
                     const uint32_t* group_rows, bool stats, size_t spill_keys) {
    const size_t vsz = f64 ? 8 : 4, G = T + 1, np = C * T;
    // job_setup: labels (32-bit and packed at <= 16 bits), group sizes, U1 and P,
    // error block, the tie term and sample sizes when kept, sort-path workspace
    size_t f = up256(R * 4) + up256(label_words(R, 16) * 4) + up256(G * 4) + 2 * up256(np * 8) +
               up256(32);
    f += 2 * up256(R * 4);   // the groups' rows and groups (matrices with columns in row order)
    if (stats) f += up256(np * 8);
    if (stats || (o.method == 1 && o.zero_handling == 0)) f += 2 * up256(np * 8);
    f += up256(nnz * vsz) + 3 * up256(C * G * 4) + up256(C * sizeof(ColPlan));
    f += up256(C * (size_t)kTabStride * 4);
    // job_upload_desc: descriptor, count-kernel items, sort-path queue
    f += up256(sizeof(JobDesc)) + up256(C * kMaxSlices * sizeof(CountItem)) +
         up256(C * sizeof(WorkItem));
    if (spill_keys) f += up256(spill_keys * vsz) + up256(C * 8);
    // exact-null tables of min/max/mix: sizes are the group row counts
    if (o.method == 1 && o.zero_handling != 0 && group_rows) {
        std::set<std::pair<int32_t, int32_t>> pairs;
        for (size_t g = 1; g <= T; ++g) pairs.emplace((int32_t)group_rows[0], (int32_t)group_rows[g]);
        size_t tab = 0, max_cells = 0;
        for (const auto& pr : pairs) {
            const size_t U = (size_t)pr.first * pr.second;
            tab += 1 + 2 * (U + 1);
            const size_t s_ = (size_t)std::min(pr.first, pr.second);
            max_cells = std::max(max_cells, (s_ + 1) * (U + 1));
        }
        f += up256(pairs.size() * sizeof(ExactTable)) + up256(tab * 8);
        f += up256(pairs.size() * 2 * max_cells * 8) + up256(T * 4);
    }
    return f;
}

size_t de_footprint(int dtype, size_t R, size_t C, size_t nnz, size_t T) {
    hpdex_mwu_opts o{};
    o.zero_handling = 1;
    o.alternative = 2;
    o.method = 2;
    const size_t G = T + 1, np = C * T;
    return matrix_footprint(dtype, R, C, nnz, HPDEX_LAYOUT_CSC, 64, false) +
           job_footprint(R, C, nnz, dtype_uses_double(dtype), T, o, nullptr, false,
                         spill_keys_bound(nullptr, C, nnz, (uint32_t)std::min(R, (size_t)UINT32_MAX),
                                          dtype_uses_double(dtype), tuning())) +
           up256(C * G * 8) + 3 * up256(np * 8) + up256(fdr_workspace_bound((int64_t)np));
}

}  // namespace

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* hpdex_version(void) { return "hpdex-mi355x 0.1 (gfx950)"; }

int hpdex_device_info(int* n_cus, int* max_lds_per_block, char* name, size_t name_len) {
    char err[256];
    if (ensure_init(err, sizeof err)) return 1;
    if (n_cus) *n_cus = g_cus;
    if (max_lds_per_block) *max_lds_per_block = g_max_lds;
    if (name && name_len) {
        std::strncpy(name, g_name, name_len - 1);
        name[name_len - 1] = 0;
    }
    return 0;
}

int hpdex_bandwidth_probe(size_t bytes, int iters, double* gbs, char* err, size_t err_len) {
    if (ensure_init(err, err_len)) return 1;
    if (bytes < (1u << 20) || iters < 1) return set_err(err, err_len, "probe needs >= 1 MiB and >= 1 iteration");
    bytes &= ~(size_t)15;
    Stream st;
    HIPCHK(st.create());
    DevBuf a, b;
    HIPCHK(a.alloc(bytes));
    HIPCHK(b.alloc(bytes));
    HIPCHK(launch_stream_copy(a.p, b.p, bytes, g_cus * 8, st.s));   // first touch
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    hipEventRecord(e0, st.s);
    for (int i = 0; i < iters; ++i) HIPCHK(launch_stream_copy(a.p, b.p, bytes, g_cus * 8, st.s));
    hipEventRecord(e1, st.s);
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (gbs) *gbs = 2.0 * (double)bytes * iters / (ms * 1e-3) / 1e9;
    return 0;
}

int hpdex_get_tuning(hpdex_tuning* out) {
    if (!out) return 1;
    const Tuning t = tuning();
    out->hist_max_bins = t.hist_max;
    out->size_ordered = t.size_ordered;
    out->count_lds_bytes = t.count_lds;
    out->count_block = t.count_block;
    out->count_packed = t.count_packed;
    out->split_arrival = t.split_arrival;
    out->ref_lds_keys = t.ref_lds_keys;
    out->ref_cache = t.ref_cache;
    return 0;
}

int hpdex_set_tuning(const hpdex_tuning* t, char* err, size_t err_len) {
    if (!t) return set_err(err, err_len, "null tuning");
    if (t->hist_max_bins < 0 || t->hist_max_bins > kHistMaxBins)
        return set_err(err, err_len, "hist_max_bins must be in [0, 8192]");
    if (t->size_ordered != 0 && t->size_ordered != 1)
        return set_err(err, err_len, "size_ordered must be 0 or 1");
    if (t->count_lds_bytes != 0 &&
        (t->count_lds_bytes < 16 * 1024 || t->count_lds_bytes > kCountLdsMax ||
         t->count_lds_bytes % 16 != 0))
        return set_err(err, err_len,
                       "count_lds_bytes must be 0 or a multiple of 16 in [16384, 155648]");
    if (t->count_block != 256 && t->count_block != 512 && t->count_block != 1024)
        return set_err(err, err_len, "count_block must be 256, 512 or 1024");
    if (t->count_packed != 0 && t->count_packed != 1)
        return set_err(err, err_len, "count_packed must be 0 or 1");
    if (t->split_arrival != 0 && t->split_arrival != 1)
        return set_err(err, err_len, "split_arrival must be 0 or 1");
    if (t->ref_lds_keys < 64 || t->ref_lds_keys > kRefKeysF32 ||
        (t->ref_lds_keys & (t->ref_lds_keys - 1)) != 0)
        return set_err(err, err_len, "ref_lds_keys must be a power of two in [64, 8192]");
    if (t->ref_cache != 0 && t->ref_cache != 1)
        return set_err(err, err_len, "ref_cache must be 0 or 1");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_init && !g_count_ready && t->count_lds_bytes != 0)
        return set_err(err, err_len, "this device does not give the count kernel its LDS");
    g_tune.hist_max = t->hist_max_bins;
    g_tune.size_ordered = t->size_ordered;
    g_tune.count_lds = t->count_lds_bytes;
    g_tune.count_block = t->count_block;
    g_tune.count_packed = t->count_packed;
    g_tune.split_arrival = t->split_arrival;
    g_tune.ref_lds_keys = t->ref_lds_keys;
    g_tune.ref_cache = t->ref_cache;
    return 0;
}

}  // extern "C"

namespace hpdex_host {

// The one-shot call on stream s (the caller checked the arguments and ran
// ensure_init): upload, the launch sequence, exact p, copy out.  Device
// memory comes from DevBuf (hipMalloc, or the calling thread's BumpBlock).
int mwu_oneshot(int dtype, const void* data, const int64_t* indices, const int64_t* indptr,
                size_t n_rows, size_t n_cols, size_t nnz, const int32_t* group_id,
                size_t n_targets, const hpdex_mwu_opts* opts, double* U1_out, double* P_out,
                double* tie_out, double* n1_out, double* n2_out, hipStream_t s, char* err,
                size_t err_len) {
    const Tuning tu = tuning();
    hpdex_matrix m;
    m.pack = false;   // used once: no packed entries, no reference tables kept
    m.resident = false;
    if (upload_matrix(&m, dtype, data, indices, indptr, n_rows, n_cols, nnz, s, err, err_len))
        return 1;
    JobState js;
    const bool stats = tie_out || n1_out || n2_out;
    if (job_setup(js, &m, group_id, n_targets, opts, stats, s, err, err_len)) return 1;
    const JobDesc J = make_desc(js, &m, tu);
    if (job_upload_desc(js, J, m, tu, false, s, err, err_len)) return 1;
    HIPCHK(launch_mwu_pipeline(pipeline_of(js, m.f64), s));
    if (job_errors(js, s, err, err_len)) return 1;
    if (opts->method == 1) {
        if (opts->zero_handling == 0) {
            if (run_exact_none(js, s, err, err_len)) return 1;
        } else {
            if (exact_plan(js, nullptr, nullptr, s, err, err_len)) return 1;
            if (exact_launch(js, s, true, err, err_len)) return 1;
        }
    }
    return copy_out(js, U1_out, P_out, tie_out, n1_out, n2_out, s, err, err_len);
}

// Arguments of the one-shot calls, checked as the reference checks them.
static int oneshot_checks(int dtype, size_t n_targets, const hpdex_mwu_opts* opts, char* err,
                          size_t err_len) {
    if (n_targets < 1)
        return set_err(err, err_len, "[mannwhitney] n_targets must be >= 1 (>=1 target groups)");
    if (!valid_dtype(dtype)) return set_err(err, err_len, "Unsupported dtype for data");
    if (!opts_ok(opts)) return set_err(err, err_len, "invalid option block");
    return 0;
}

}  // namespace hpdex_host

extern "C" {

int hpdex_mwu_csc_stats(int dtype, const void* data, const int64_t* indices,
                        const int64_t* indptr, size_t n_rows, size_t n_cols, size_t nnz,
                        const int32_t* group_id, size_t n_targets, const hpdex_mwu_opts* opts,
                        int threads, double* U1_out, double* P_out, double* tie_out,
                        double* n1_out, double* n2_out, char* err, size_t err_len) {
    (void)threads;
    if (oneshot_checks(dtype, n_targets, opts, err, err_len)) return 1;
    if (n_cols == 0) return 0;
    if (ensure_init(err, err_len)) return 1;
    Stream st;
    HIPCHK(st.create());
    return mwu_oneshot(dtype, data, indices, indptr, n_rows, n_cols, nnz, group_id, n_targets,
                       opts, U1_out, P_out, tie_out, n1_out, n2_out, st.s, err, err_len);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// hpdex_mwu_csc_noalloc: the one-shot call with every device buffer carved
// from this file-scope array instead of hipMalloc.  Each GPU holds its own
// copy of the array, at its own address, until the process ends.  In host
// code `scratch` names a host-side stand-in: its device address comes from
// hipGetSymbolAddress, once per device, and only that address is used.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kNoAllocMaxDevices = 64;
struct NoAllocDevice {
    void* ptr = nullptr;        // device address of `scratch` on this device
    hipStream_t s = nullptr;    // the stream its calls run on
    bool ready = false;
};
// Calls share their device's scratch: the lock runs them one at a time.
std::mutex g_noalloc_mu;
NoAllocDevice g_noalloc[kNoAllocMaxDevices];

// Bind the calling thread's DevBuf allocations to a block for one scope.
struct BlockScope {
    explicit BlockScope(BumpBlock* b) { t_block = b; }
    ~BlockScope() { t_block = nullptr; }
};

}  // namespace

extern "C" {

int hpdex_mwu_csc_noalloc(int dtype, const void* data, const int64_t* indices,
                          const int64_t* indptr, size_t n_rows, size_t n_cols, size_t nnz,
                          const int32_t* group_id, size_t n_targets, const hpdex_mwu_opts* opts,
                          int threads, double* U1_out, double* P_out, char* err,
                          size_t err_len) {
    (void)threads;
    if (oneshot_checks(dtype, n_targets, opts, err, err_len)) return 1;
    if (n_cols == 0) return 0;
    if (ensure_init(err, err_len)) return 1;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    // host copies without pinned staging (its buffers are allocations too)
    struct PlainCopies {
        PlainCopies() { t_plain_copies = true; }
        ~PlainCopies() { t_plain_copies = false; }
    } plain_copies;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices)
        return set_err(err, err_len, "hpdex_mwu_csc_noalloc: device index out of range");
    NoAllocDevice& d = g_noalloc[dev];
    if (!d.ready) {   // set-up on this device's first call, with the device current
        HIPCHK(hipGetSymbolAddress(&d.ptr, HIP_SYMBOL(scratch)));
        HIPCHK(hipStreamCreateWithFlags(&d.s, hipStreamNonBlocking));
        d.ready = true;
    }
    // everything the call carves from the scratch, before anything runs
    if (n_rows > (size_t)INT32_MAX) return set_err(err, err_len, "more than 2^31-1 rows is unsupported");
    std::vector<uint32_t> group_rows(n_targets + 1, 0u);
    for (size_t r = 0; r < n_rows; ++r) {
        const int g = group_id[r];
        if (g >= 0 && (size_t)g <= n_targets) ++group_rows[g];
    }
    const size_t need =
        matrix_footprint(dtype, n_rows, n_cols, nnz, HPDEX_LAYOUT_CSC, 64, false) +
        job_footprint(n_rows, n_cols, nnz, dtype_uses_double(dtype), n_targets, *opts,
                      group_rows.data(), false,
                      spill_keys_bound(indptr, n_cols, nnz, group_rows[0],
                                       dtype_uses_double(dtype), tuning()));
    if (need > kNoAllocScratchBytes)
        return set_err(err, err_len,
                       "hpdex_mwu_csc_noalloc: the job needs " + std::to_string(need) +
                           " bytes of device memory, more than its scratch of " +
                           std::to_string(kNoAllocScratchBytes) + " bytes; use hpdex_mwu_csc");
    BumpBlock blk;
    blk.base = (char*)d.ptr;
    blk.cap = kNoAllocScratchBytes;
    int rc;
    {
        BlockScope scope(&blk);
        rc = mwu_oneshot(dtype, data, indices, indptr, n_rows, n_cols, nnz, group_id, n_targets,
                         opts, U1_out, P_out, nullptr, nullptr, nullptr, d.s, err, err_len);
    }
    // nothing of this call may still run when the next one reuses the scratch
    const hipError_t se = hipStreamSynchronize(d.s);
    if (rc && blk.exhausted)   // exact p in mode none: tables sized by the data
        return set_err(err, err_len,
                       "hpdex_mwu_csc_noalloc: the job's exact-null tables do not fit its "
                       "scratch of " + std::to_string(kNoAllocScratchBytes) +
                           " bytes; use hpdex_mwu_csc");
    if (rc) return rc;
    if (se != hipSuccess) return set_err(err, err_len, std::string("HIP error: ") + hipGetErrorString(se));
    return 0;
}

int hpdex_mwu_csc(int dtype, const void* data, const int64_t* indices, const int64_t* indptr,
                  size_t n_rows, size_t n_cols, size_t nnz, const int32_t* group_id,
                  size_t n_targets, const hpdex_mwu_opts* opts, int threads, double* U1_out,
                  double* P_out, char* err, size_t err_len) {
    return hpdex_mwu_csc_stats(dtype, data, indices, indptr, n_rows, n_cols, nnz, group_id,
                               n_targets, opts, threads, U1_out, P_out, nullptr, nullptr,
                               nullptr, err, err_len);
}

int hpdex_group_mean_csc(int dtype, const void* data, const int64_t* indices,
                         const int64_t* indptr, size_t n_rows, size_t n_cols, size_t nnz,
                         const int32_t* group_id, size_t n_groups, int include_zeros,
                         int use_kahan, int threads, double* out, char* err, size_t err_len) {
    (void)threads;
    (void)use_kahan;   // the device sums in FP64 either way
    if (!valid_dtype(dtype)) return set_err(err, err_len, "Unsupported dtype for data");
    if (n_groups == 0 || n_cols == 0) return 0;
    if (ensure_init(err, err_len)) return 1;
    const Tuning tu = tuning();
    Stream st;
    HIPCHK(st.create());
    hpdex_matrix m;
    m.pack = false;   // used once: no packed entries, no reference tables kept
    m.resident = false;

    if (upload_matrix(&m, dtype, data, indices, indptr, n_rows, n_cols, nnz, st.s, err, err_len))
        return 1;
    std::vector<uint32_t> grows(n_groups, 0u);
    for (size_t r = 0; r < n_rows; ++r) {
        const int g = group_id[r];
        if (g >= 0 && (size_t)g < n_groups) ++grows[g];
    }
    DevBuf gid, dg, dout;
    HIPCHK(gid.alloc(n_rows * 4));
    if (n_rows) HIPCHK(hipMemcpyAsync(gid.p, group_id, n_rows * 4, hipMemcpyHostToDevice, st.s));
    HIPCHK(dg.alloc(n_groups * 4));
    HIPCHK(hipMemcpyAsync(dg.p, grows.data(), n_groups * 4, hipMemcpyHostToDevice, st.s));
    HIPCHK(dout.alloc(n_cols * n_groups * 8));
    HIPCHK(launch_group_mean(m.vals.p, m.f64, m.rows.as<int32_t>(), m.indptr.as<int64_t>(),
                             gid.as<int32_t>(), dg.as<uint32_t>(), (int64_t)n_cols,
                             (int32_t)n_rows, (int32_t)n_groups, include_zeros, dout.as<double>(),
                             st.s));
    HIPCHK(hipMemcpyAsync(out, dout.p, n_cols * n_groups * 8, hipMemcpyDeviceToHost, st.s));
    HIPCHK(hipStreamSynchronize(st.s));
    return 0;
}

int hpdex_fdr_bh(const double* p, size_t n, double* fdr_out, char* err, size_t err_len) {
    if (n == 0) return 0;
    if (!p || !fdr_out) return set_err(err, err_len, "null argument");
    if (n > (size_t)UINT32_MAX) return set_err(err, err_len, "more than 2^32-1 p-values");
    if (ensure_init(err, err_len)) return 1;
    Stream st;
    HIPCHK(st.create());
    DevBuf dp, df, ws;
    HIPCHK(dp.alloc(n * 8));
    HIPCHK(df.alloc(n * 8));
    const size_t wsb = fdr_workspace_bytes((int64_t)n);
    if (wsb > fdr_workspace_bound((int64_t)n))
        return set_err(err, err_len, "FDR workspace above its stated bound");
    HIPCHK(ws.alloc(wsb));
    HIPCHK(hipMemcpyAsync(dp.p, p, n * 8, hipMemcpyHostToDevice, st.s));
    HIPCHK(launch_fdr_bh(dp.as<double>(), (int64_t)n, df.as<double>(), ws.p, wsb, st.s));
    HIPCHK(hipMemcpyAsync(fdr_out, df.p, n * 8, hipMemcpyDeviceToHost, st.s));
    HIPCHK(hipStreamSynchronize(st.s));
    return 0;
}

// de.py:120-184 on a matrix already on the device (steps 4-7 of the driver)
static int de_run(hpdex_matrix& m, const int32_t* group_id, size_t n_targets, int tie_correction,
                  int use_continuity, double clip_value, double* U1_out, double* P_out,
                  double* fc_out, double* log2fc_out, double* fdr_out, hipStream_t s, char* err,
                  size_t err_len) {
    const size_t n_cols = m.C, n_rows = m.R;
    const size_t G = n_targets + 1, np = n_cols * n_targets;
    const Tuning tu = tuning();
    // de.py:147-158: zero_handling "min", two-sided, asymptotic
    hpdex_mwu_opts o{};
    o.tie_correction = tie_correction ? 1 : 0;
    o.use_continuity = use_continuity ? 1 : 0;
    o.zero_handling = 1;
    o.alternative = 2;
    o.method = 2;
    JobState js;
    if (job_setup(js, &m, group_id, n_targets, &o, false, s, err, err_len)) return 1;
    const JobDesc J = make_desc(js, &m, tu);
    if (job_upload_desc(js, J, m, tu, false, s, err, err_len)) return 1;
    // de.py:126-142: group means with zeros, fold change, log2 fold change
    DevBuf means, fc, lfc, fdr, ws;
    HIPCHK(means.alloc(n_cols * G * 8));
    HIPCHK(launch_group_mean(m.vals.p, m.f64, m.rows.as<int32_t>(), m.indptr.as<int64_t>(),
                             js.gid.as<int32_t>(), js.grows.as<uint32_t>(), (int64_t)n_cols,
                             (int32_t)n_rows, (int32_t)G, 1, means.as<double>(), s));
    HIPCHK(fc.alloc(np * 8));
    HIPCHK(lfc.alloc(np * 8));
    HIPCHK(launch_fold_change(means.as<double>(), (int64_t)n_cols, (int32_t)n_targets, clip_value,
                              fc.as<double>(), lfc.as<double>(), s));
    // de.py:147-158: the test
    HIPCHK(launch_mwu_pipeline(pipeline_of(js, m.f64), s));
    if (job_errors(js, s, err, err_len)) return 1;
    // de.py:172-184: BH FDR over every p-value
    HIPCHK(fdr.alloc(np * 8));
    const size_t wsb = fdr_workspace_bytes((int64_t)np);
    // de_footprint counts the bound (it is computed without the device)
    if (wsb > fdr_workspace_bound((int64_t)np))
        return set_err(err, err_len, "FDR workspace above its stated bound");
    HIPCHK(ws.alloc(wsb));
    HIPCHK(launch_fdr_bh(js.P.as<double>(), (int64_t)np, fdr.as<double>(), ws.p, wsb, s));
    const size_t b = np * 8;
    if (U1_out) HIPCHK(download_bytes(U1_out, js.U1.p, b, s));
    if (P_out) HIPCHK(download_bytes(P_out, js.P.p, b, s));
    if (fc_out) HIPCHK(download_bytes(fc_out, fc.p, b, s));
    if (log2fc_out) HIPCHK(download_bytes(log2fc_out, lfc.p, b, s));
    if (fdr_out) HIPCHK(download_bytes(fdr_out, fdr.p, b, s));
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

static int de_checks(size_t n_targets, size_t n_cols, char* err, size_t err_len) {
    if (n_targets < 1)
        return set_err(err, err_len, "[mannwhitney] n_targets must be >= 1 (>=1 target groups)");
    if (n_cols * n_targets > (size_t)UINT32_MAX) return set_err(err, err_len, "more than 2^32-1 tests");
    return 0;
}

int hpdex_de_csc(int dtype, const void* data, const int64_t* indices, const int64_t* indptr,
                 size_t n_rows, size_t n_cols, size_t nnz, const int32_t* group_id,
                 size_t n_targets, int tie_correction, int use_continuity, double clip_value,
                 int threads, double* U1_out, double* P_out, double* fc_out,
                 double* log2fc_out, double* fdr_out, char* err, size_t err_len) {
    (void)threads;
    if (de_checks(n_targets, n_cols, err, err_len)) return 1;
    if (!valid_dtype(dtype)) return set_err(err, err_len, "Unsupported dtype for data");
    if (n_cols == 0) return 0;
    if (ensure_init(err, err_len)) return 1;
    Stream st;
    HIPCHK(st.create());
    hpdex_matrix m;
    m.pack = false;   // used once: no packed entries, no reference tables kept
    m.resident = false;

    if (upload_matrix(&m, dtype, data, indices, indptr, n_rows, n_cols, nnz, st.s, err, err_len))
        return 1;
    return de_run(m, group_id, n_targets, tie_correction, use_continuity, clip_value, U1_out, P_out,
                  fc_out, log2fc_out, fdr_out, st.s, err, err_len);
}

int hpdex_de_matrix(hpdex_matrix* m, const int32_t* group_id, size_t n_targets,
                    int tie_correction, int use_continuity, double clip_value, double* U1_out,
                    double* P_out, double* fc_out, double* log2fc_out, double* fdr_out, char* err,
                    size_t err_len) {
    if (!m) return set_err(err, err_len, "null handle");
    if (de_checks(n_targets, m->C, err, err_len)) return 1;
    if (m->C == 0) return 0;
    if (ensure_init(err, err_len)) return 1;
    Stream st;
    HIPCHK(st.create());
    return de_run(*m, group_id, n_targets, tie_correction, use_continuity, clip_value, U1_out, P_out,
                  fc_out, log2fc_out, fdr_out, st.s, err, err_len);
}

// The numba pipeline (src/hpdex/backend/de_.py:508-658 with the kernels of
// pykernel_.py) on a resident matrix: the Mann-Whitney pass under "mix"
// (the pipeline densifies X, so zeros rank at their true position and n is
// the group's row count) with the raw tie sums kept and no small-sample
// check; group sums for the fold change; the per-(target, gene chunk) method
// choice; the exact tables the chosen chunks need; target U and p, FDR.
static int pden_run(hpdex_matrix& m, const int32_t* group_id, size_t T, const hpdex_pden_opts& po,
                    double* U_out, double* P_out, double* fc_out, double* log2fc_out,
                    double* fdr_out, hipStream_t s, char* err, size_t err_len) {
    const size_t C = m.C, R = m.R, G = T + 1, np = C * T;
    const Tuning tu = tuning();
    hpdex_mwu_opts o{};
    o.tie_correction = 1;   // the raw tie sums: the p step applies tie_correction itself
    o.use_continuity = po.continuity_correction ? 1 : 0;
    o.zero_handling = 3;
    o.alternative = 2;
    o.method = 1;           // U and ties only: the pass writes no p
    JobState js;
    if (job_setup(js, &m, group_id, T, &o, true, s, err, err_len)) return 1;
    JobDesc J = make_desc(js, &m, tu);
    J.asym = 0;
    J.min_n = 0u;
    if (job_upload_desc(js, J, m, tu, false, s, err, err_len)) return 1;
    DevBuf sums, fc, lfc;
    HIPCHK(sums.alloc(C * G * 8));
    HIPCHK(launch_pden_group_sums(m.vals.p, m.f64, m.dtype == HPDEX_FLOAT32,
                                  m.rows.as<int32_t>(), m.indptr.as<int64_t>(),
                                  js.gid.as<int32_t>(), (int64_t)C, (int32_t)R, (int32_t)G,
                                  sums.as<double>(), s));
    HIPCHK(fc.alloc(np * 8));
    HIPCHK(lfc.alloc(np * 8));
    HIPCHK(launch_pden_fold_change(sums.as<double>(), js.grows.as<uint32_t>(), (int64_t)C,
                                   (int32_t)T, po.clip_value, m.dtype == HPDEX_FLOAT32,
                                   fc.as<double>(), lfc.as<double>(), s));
    HIPCHK(launch_mwu_pipeline(pipeline_of(js, m.f64), s));
    if (job_errors(js, s, err, err_len)) return 1;
    // the method of each (target, gene chunk)
    const int64_t chunk = po.chunk_genes;
    const int64_t nchunk = ((int64_t)C + chunk - 1) / chunk;
    DevBuf kmin, kmax, tflag, dexact;
    const bool hist = po.use_hist != 0;
    HIPCHK(kmin.alloc(hist ? (size_t)nchunk * G * 8 : 16));
    HIPCHK(kmax.alloc(hist ? (size_t)nchunk * G * 8 : 16));
    HIPCHK(tflag.alloc((size_t)nchunk * T * 4));
    HIPCHK(dexact.alloc((size_t)nchunk * T));
    HIPCHK(launch_pden_decide(m.vals.p, m.f64, m.rows.as<int32_t>(), m.indptr.as<int64_t>(),
                              js.gid.as<int32_t>(), js.grows.as<uint32_t>(), (int64_t)C, (int32_t)R,
                              (int32_t)T, js.tie.as<double>(), chunk, po.use_asymptotic, hist,
                              (double)po.max_bins, kmin.as<unsigned long long>(),
                              kmax.as<unsigned long long>(), tflag.as<uint32_t>(),
                              dexact.as<uint8_t>(), s));
    std::vector<uint8_t> ex((size_t)nchunk * T);
    HIPCHK(hipMemcpyAsync(ex.data(), dexact.p, ex.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    // exact tables of the targets with an exact chunk, one per distinct size
    std::vector<ExactTable> tabs;
    std::vector<int32_t> tidx(T, -1);
    std::map<int32_t, int32_t> by_n;
    int64_t off = 0;
    size_t max_cells = 0;
    for (size_t t = 0; t < T; ++t) {
        bool any = false;
        for (int64_t k = 0; k < nchunk && !any; ++k) any = ex[t * nchunk + k] != 0;
        if (!any) continue;
        const int32_t nt = (int32_t)js.group_rows[t + 1], nr = (int32_t)js.group_rows[0];
        auto it = by_n.find(nt);
        if (it != by_n.end()) {
            tidx[t] = it->second;
            continue;
        }
        const int64_t U = (int64_t)nr * nt;
        const size_t cells = ((size_t)std::min(nr, nt) + 1) * (size_t)(U + 1);
        if (cells > kExactMaxCells)
            return set_err(err, err_len,
                           "exact null table too large for n_ref=" + std::to_string(nr) +
                               ", n_tar=" + std::to_string(nt) + "; use use_asymptotic=True");
        max_cells = std::max(max_cells, cells);
        ExactTable tb;
        tb.offset = off;
        tb.n1 = nr;
        tb.n2 = nt;
        off += 1 + 2 * (U + 1);
        tidx[t] = (int32_t)tabs.size();
        by_n[nt] = tidx[t];
        tabs.push_back(tb);
    }
    DevBuf d_tabs, tabbuf, tscratch, d_tidx;
    HIPCHK(d_tidx.alloc(T * 4));
    HIPCHK(hipMemcpyAsync(d_tidx.p, tidx.data(), T * 4, hipMemcpyHostToDevice, s));
    if (!tabs.empty()) {
        HIPCHK(d_tabs.alloc(tabs.size() * sizeof(ExactTable)));
        HIPCHK(hipMemcpyAsync(d_tabs.p, tabs.data(), tabs.size() * sizeof(ExactTable),
                              hipMemcpyHostToDevice, s));
        HIPCHK(tabbuf.alloc((size_t)off * 8));
        HIPCHK(tscratch.alloc(tabs.size() * 2 * max_cells * 8));
        HIPCHK(launch_exact_tables(d_tabs.as<ExactTable>(), (int)tabs.size(), tabbuf.as<double>(),
                                   tscratch.as<double>(), 2 * max_cells, s));
    }
    DevBuf U, P, fdr, ws;
    HIPCHK(U.alloc(np * 8));
    HIPCHK(P.alloc(np * 8));
    PdenJob PJ{};
    PJ.U1 = js.U1.as<double>();
    PJ.tie = js.tie.as<double>();
    PJ.group_rows = js.grows.as<uint32_t>();
    PJ.exact = dexact.as<uint8_t>();
    PJ.tidx = d_tidx.as<int32_t>();
    PJ.tables = d_tabs.as<ExactTable>();
    PJ.tabbuf = tabbuf.as<double>();
    PJ.U_out = U.as<double>();
    PJ.P_out = P.as<double>();
    PJ.C = (int64_t)C;
    PJ.chunk = chunk;
    PJ.nchunk = nchunk;
    PJ.T = (int32_t)T;
    PJ.tie_corr = po.tie_correction ? 1 : 0;
    PJ.cc = po.continuity_correction ? 0.5 : 0.0;
    HIPCHK(launch_pden_finish(PJ, s));
    // BH over every p-value (de_.py:655-656, _compute_fdr)
    HIPCHK(fdr.alloc(np * 8));
    const size_t wsb = fdr_workspace_bytes((int64_t)np);
    HIPCHK(ws.alloc(wsb ? wsb : 16));
    HIPCHK(launch_fdr_bh(P.as<double>(), (int64_t)np, fdr.as<double>(), ws.p, wsb, s));
    const size_t b = np * 8;
    if (U_out) HIPCHK(download_bytes(U_out, U.p, b, s));
    if (P_out) HIPCHK(download_bytes(P_out, P.p, b, s));
    if (fc_out) HIPCHK(download_bytes(fc_out, fc.p, b, s));
    if (log2fc_out) HIPCHK(download_bytes(log2fc_out, lfc.p, b, s));
    if (fdr_out) HIPCHK(download_bytes(fdr_out, fdr.p, b, s));
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

int hpdex_pden_matrix(hpdex_matrix* m, const int32_t* group_id, size_t n_targets,
                      const hpdex_pden_opts* o, double* U_out, double* P_out, double* fc_out,
                      double* log2fc_out, double* fdr_out, char* err, size_t err_len) {
    if (!m || !o) return set_err(err, err_len, "null argument");
    if (de_checks(n_targets, m->C, err, err_len)) return 1;
    if (o->use_asymptotic < -1 || o->use_asymptotic > 1)
        return set_err(err, err_len, "use_asymptotic must be -1 (None), 0 or 1");
    if (o->chunk_genes < 1) return set_err(err, err_len, "chunk_genes must be at least 1");
    if (m->C == 0) return 0;
    if (n_targets * m->C > (size_t)UINT32_MAX)
        return set_err(err, err_len, "more than 2^32-1 (gene, target) pairs");
    if (ensure_init(err, err_len)) return 1;
    Stream st;
    HIPCHK(st.create());
    return pden_run(*m, group_id, n_targets, *o, U_out, P_out, fc_out, log2fc_out, fdr_out, st.s,
                    err, err_len);
}

int hpdex_matrix_create_from(const hpdex_matrix_src* src, hpdex_matrix** out, char* err,
                             size_t err_len) {
    if (!out || !src) return set_err(err, err_len, "null argument");
    *out = nullptr;
    if (ensure_init(err, err_len)) return 1;
    Stream st;
    HIPCHK(st.create());
    hpdex_matrix* m = new hpdex_matrix();
    if (upload_matrix_from(m, src, st.s, err, err_len)) {
        delete m;
        return 1;
    }
    *out = m;
    return 0;
}

int hpdex_matrix_shape(const hpdex_matrix* m, size_t* n_rows, size_t* n_cols, size_t* nnz) {
    if (!m) return 1;
    if (n_rows) *n_rows = m->R;
    if (n_cols) *n_cols = m->C;
    if (nnz) *nnz = m->nnz;
    return 0;
}

int hpdex_matrix_download(const hpdex_matrix* m, double* data_out, int32_t* rows_out,
                          int64_t* indptr_out, char* err, size_t err_len) {
    if (!m) return set_err(err, err_len, "null handle");
    if (ensure_init(err, err_len)) return 1;
    Stream st;
    HIPCHK(st.create());
    const size_t nnz = m->nnz;
    if (rows_out && nnz)
        HIPCHK(hipMemcpyAsync(rows_out, m->rows.p, nnz * 4, hipMemcpyDeviceToHost, st.s));
    if (indptr_out)
        HIPCHK(hipMemcpyAsync(indptr_out, m->indptr.p, (m->C + 1) * 8, hipMemcpyDeviceToHost, st.s));
    if (data_out && nnz) {
        if (m->f64) {
            HIPCHK(hipMemcpyAsync(data_out, m->vals.p, nnz * 8, hipMemcpyDeviceToHost, st.s));
        } else {
            std::vector<float> f(nnz);
            HIPCHK(hipMemcpyAsync(f.data(), m->vals.p, nnz * 4, hipMemcpyDeviceToHost, st.s));
            HIPCHK(hipStreamSynchronize(st.s));
            for (size_t i = 0; i < nnz; ++i) data_out[i] = (double)f[i];
        }
    }
    HIPCHK(hipStreamSynchronize(st.s));
    return 0;
}

int hpdex_matrix_create(int dtype, const void* data, const int64_t* indices,
                        const int64_t* indptr, size_t n_rows, size_t n_cols, size_t nnz,
                        hpdex_matrix** out, char* err, size_t err_len) {
    if (!out) return set_err(err, err_len, "null output handle");
    *out = nullptr;
    if (ensure_init(err, err_len)) return 1;
    Stream st;
    HIPCHK(st.create());
    hpdex_matrix* m = new hpdex_matrix();
    if (upload_matrix(m, dtype, data, indices, indptr, n_rows, n_cols, nnz, st.s, err, err_len)) {
        delete m;
        return 1;
    }
    *out = m;
    return 0;
}

void hpdex_matrix_destroy(hpdex_matrix* m) { delete m; }

size_t hpdex_de_csc_device_bytes(int dtype, size_t n_rows, size_t n_cols, size_t nnz,
                                 size_t n_targets) {
    if (!valid_dtype(dtype) || n_targets < 1) return 0;
    return de_footprint(dtype, n_rows, n_cols, nnz, n_targets);
}

int hpdex_matrix_upload_form(const hpdex_matrix* m, int64_t* patched_columns) {
    if (patched_columns) *patched_columns = m && m->compact ? m->compact_patched : 0;
    return m && m->compact ? 1 : 0;
}

int hpdex_matrix_ref_cache_info(const hpdex_matrix* m, int64_t* columns, size_t* bytes) {
    if (!m) return 1;
    std::shared_ptr<RefCacheEntry> e;
    {
        std::lock_guard<std::mutex> lk(m->rc_mu);
        e = m->rc;
    }
    if (columns) *columns = e ? e->n_cols : 0;
    if (bytes) *bytes = e ? e->bytes : 0;
    return 0;
}

void hpdex_matrix_ref_cache_clear(hpdex_matrix* m) {
    if (!m) return;
    std::lock_guard<std::mutex> lk(m->rc_mu);
    m->rc.reset();
}

size_t hpdex_matrix_device_bytes(const hpdex_matrix* m) {
    if (!m) return 0;
    return m->vals.n + m->rows.n + m->indptr.n + m->colrange.n + m->colident.n + m->colint.n +
           m->packed.n;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Plans (resident matrix, plain or batched form)
// ---------------------------------------------------------------------------
struct hpdex_plan {
    hpdex_matrix* m = nullptr;
    int form = 1;
    Tuning tu;
    Stream st;
    JobState js;
    JobDesc J{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~hpdex_plan() {
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
    }
};

namespace hpdex_host {

// The per-job launch sequence (everything after upload): split, reference
// tables, targets and pairs, then the exact p lookup when asked for (tables
// of min/max/mix depend only on group sizes and are built once per plan).
int plan_enqueue(hpdex_plan* p, char* err, size_t err_len) {
    const hipStream_t s = p->st.s;
    HIPCHK(launch_mwu_pipeline(pipeline_of(p->js, p->m->f64), s));
    if (p->js.opts.method == 1 && p->js.opts.zero_handling != 0)
        if (exact_launch(p->js, s, false, err, err_len)) return 1;
    return 0;
}

}  // namespace

extern "C" {

int hpdex_plan_create(hpdex_matrix* m, const int32_t* group_id, size_t n_targets,
                      const hpdex_mwu_opts* opts, int form, hpdex_plan** out, char* err,
                      size_t err_len) {
    if (!out || !m) return set_err(err, err_len, "null handle");
    *out = nullptr;
    if (!opts_ok(opts)) return set_err(err, err_len, "invalid option block");
    if (form < 0 || form > (HPDEX_PLAN_BATCHED | HPDEX_PLAN_STATS))
        return set_err(err, err_len,
                       "form must be HPDEX_PLAN_PLAIN or HPDEX_PLAN_BATCHED, optionally with "
                       "HPDEX_PLAN_STATS");
    const bool stats = (form & HPDEX_PLAN_STATS) != 0;
    form &= HPDEX_PLAN_BATCHED;
    if (ensure_init(err, err_len)) return 1;
    hpdex_plan* p = new hpdex_plan();
    p->m = m;
    p->form = form;
    p->tu = tuning();
    auto bail = [&](int rc) { delete p; return rc; };
    if (p->st.create() != hipSuccess) return bail(set_err(err, err_len, "stream creation failed"));
    if (job_setup(p->js, m, group_id, n_targets, opts, stats, p->st.s, err, err_len))
        return bail(1);
    p->J = make_desc(p->js, m, p->tu);
    const bool ordered = form == 1 && p->tu.size_ordered;
    if (job_upload_desc(p->js, p->J, *m, p->tu, ordered, p->st.s, err, err_len)) return bail(1);
    if (opts->method == 1 && opts->zero_handling != 0) {
        if (exact_plan(p->js, nullptr, nullptr, p->st.s, err, err_len)) return bail(1);
        const hipError_t e = launch_exact_tables(p->js.d_tabs.as<ExactTable>(),
                                                 (int)p->js.tabs.size(), p->js.tabbuf.as<double>(),
                                                 p->js.tscratch.as<double>(),
                                                 p->js.scratch_per_table, p->st.s);
        if (e != hipSuccess) return bail(set_err(err, err_len, hipGetErrorString(e)));
    }
    if (hipStreamSynchronize(p->st.s) != hipSuccess)
        return bail(set_err(err, err_len, "plan setup failed"));
    if (form == 1 && m->C > 0) {
        if (hipStreamBeginCapture(p->st.s, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return bail(set_err(err, err_len, "graph capture failed to start"));
        const int rc = plan_enqueue(p, err, err_len);
        const hipError_t e = hipStreamEndCapture(p->st.s, &p->graph);
        if (rc) return bail(1);
        if (e != hipSuccess) return bail(set_err(err, err_len, hipGetErrorString(e)));
        if (hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0) != hipSuccess)
            return bail(set_err(err, err_len, "graph instantiation failed"));
    }
    *out = p;
    return 0;
}

void hpdex_plan_destroy(hpdex_plan* p) { delete p; }

int hpdex_plan_get_info(hpdex_plan* p, hpdex_plan_info* out, char* err, size_t err_len) {
    if (!p || !out) return set_err(err, err_len, "null handle");
    out->count_columns = p->js.n_count_cols;
    out->count_blocks = p->js.n_count;
    out->count_packed_columns = p->js.n_count_packed;
    out->count_nnz = p->js.count_nnz;
    out->split_columns = p->js.n_gen_static;
    out->handed_over = 0;   // the count kernel reports a column it cannot take as an error
    out->ref_cached_columns = p->js.rc_cols;
    return 0;
}

int hpdex_plan_stage_times(hpdex_plan* p, int iters, double* ms, char* err, size_t err_len) {
    if (!p || !ms) return set_err(err, err_len, "null handle");
    if (iters < 1) iters = 1;
    for (int k = 0; k < 4; ++k) ms[k] = 0.0;
    if (p->m->C == 0) return 0;
    const hipStream_t s = p->st.s;
    std::vector<hipEvent_t> ev(5 * (size_t)iters, nullptr);
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    int rc = 0;
    for (int it = 0; it < iters && !rc; ++it) {
        const hipError_t e = launch_mwu_pipeline(pipeline_of(p->js, p->m->f64), s, &ev[5 * it]);
        if (e != hipSuccess) rc = set_err(err, err_len, hipGetErrorString(e));
        if (!rc && p->js.opts.method == 1 && p->js.opts.zero_handling != 0)
            rc = exact_launch(p->js, s, false, err, err_len);
    }
    const hipError_t se = hipStreamSynchronize(s);
    if (!rc && se != hipSuccess) rc = set_err(err, err_len, hipGetErrorString(se));
    if (!rc) {
        for (int it = 0; it < iters; ++it)
            for (int k = 0; k < 4; ++k) {
                float x = 0.f;
                hipEventElapsedTime(&x, ev[5 * it + k], ev[5 * it + k + 1]);
                ms[k] += x / iters;
            }
    }
    for (auto& e : ev) hipEventDestroy(e);
    return rc;
}

int hpdex_plan_kernel_ms(hpdex_plan* p, int kernel, int iters, double* ms, char* err,
                         size_t err_len) {
    if (!p || !ms) return set_err(err, err_len, "null handle");
    if (kernel < 0 || kernel > 3) return set_err(err, err_len, "kernel must be 0, 1, 2 or 3");
    if (iters < 1) iters = 1;
    *ms = 0.0;
    if (p->m->C == 0) return 0;
    const hipStream_t s = p->st.s;
    const PipelineLaunch L = pipeline_of(p->js, p->m->f64);
    // the whole sequence first, so that every stage finds its inputs
    HIPCHK(launch_mwu_pipeline(L, s));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int rc = 0;
    hipError_t e = hipEventRecord(e0, s);
    for (int it = 0; it < iters && e == hipSuccess; ++it) e = launch_mwu_stage(L, kernel, s);
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    // the whole sequence again: a split run alone leaves the reference keys
    // unsorted until the reference kernel follows
    if (e == hipSuccess) e = launch_mwu_pipeline(L, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        rc = set_err(err, err_len, hipGetErrorString(e));
    } else {
        float x = 0.f;
        hipEventElapsedTime(&x, e0, e1);
        *ms = (double)x / iters;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return rc;
}

int hpdex_plan_run(hpdex_plan* p, int iters, double* ms_total, double* kernel_ms, char* err,
                   size_t err_len) {
    if (!p) return set_err(err, err_len, "null plan");
    if (iters < 1) iters = 1;
    if (p->m->C == 0) {
        if (ms_total) *ms_total = 0;
        if (kernel_ms) *kernel_ms = 0;
        return 0;
    }
    const hipStream_t s = p->st.s;
    std::vector<hipEvent_t> ev(2 * (size_t)iters, nullptr);
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    int rc = 0;
    for (int it = 0; it < iters && !rc; ++it) {
        hipEventRecord(ev[2 * it], s);
        if (p->form == 1) {
            const hipError_t e = hipGraphLaunch(p->exec, s);
            if (e != hipSuccess) rc = set_err(err, err_len, hipGetErrorString(e));
        } else {
            rc = plan_enqueue(p, err, err_len);
        }
        hipEventRecord(ev[2 * it + 1], s);
    }
    const hipError_t se = hipStreamSynchronize(s);
    if (!rc && se != hipSuccess) rc = set_err(err, err_len, hipGetErrorString(se));
    if (!rc) {
        float tot = 0.f, sum = 0.f;
        hipEventElapsedTime(&tot, ev[0], ev[2 * iters - 1]);
        for (int it = 0; it < iters; ++it) {
            float x = 0.f;
            hipEventElapsedTime(&x, ev[2 * it], ev[2 * it + 1]);
            sum += x;
        }
        if (ms_total) *ms_total = tot;
        if (kernel_ms) *kernel_ms = sum / iters;
    }
    for (auto& e : ev) hipEventDestroy(e);
    if (!rc && p->js.opts.method == 1 && p->js.opts.zero_handling == 0) {
        if (job_errors(p->js, s, err, err_len)) return 1;
        rc = run_exact_none(p->js, s, err, err_len);
    }
    // the matrix keeps this job's sorted reference after its first run
    if (!rc) rc = refcache_fill(p->js, *p->m, s, err, err_len);
    return rc;
}

int hpdex_plan_results(hpdex_plan* p, double* U1_out, double* P_out, double* tie_out,
                       double* n1_out, double* n2_out, char* err, size_t err_len) {
    if (!p) return set_err(err, err_len, "null plan");
    if (job_errors(p->js, p->st.s, err, err_len)) return 1;
    return copy_out(p->js, U1_out, P_out, tie_out, n1_out, n2_out, p->st.s, err, err_len);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches: several plans run as one launch sequence
// ---------------------------------------------------------------------------
struct hpdex_batch {
    std::vector<hpdex_plan*> plans;
    Stream st;
    DevBuf d_jobs[2], d_count[2], d_gen[2];   // [0] float matrices, [1] double
    BatchLaunch L[2]{};
    int64_t n_count_items = 0, n_gen_items = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~hpdex_batch() {
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
    }
};

namespace hpdex_host {

// Merge per-job lists, each ordered by a non-increasing key, into one list
// ordered the same way (ties: the lower job first) with a heap over the
// jobs' cursors: O(n log jobs), where sorting the concatenation is
// O(n log n) (200,000 items: 21 ms on one host core).  A job's list that is
// not ordered (a plan laid out in column order) is sorted first.
template <typename T, typename Key>
static void merge_ordered(std::vector<std::vector<T>>& lists, Key key, std::vector<T>& out) {
    size_t total = 0;
    for (auto& l : lists) {
        total += l.size();
        bool ordered = true;
        for (size_t k = 1; k < l.size() && ordered; ++k) ordered = key(l[k - 1]) >= key(l[k]);
        if (!ordered)
            std::stable_sort(l.begin(), l.end(),
                             [&](const T& a, const T& c) { return key(a) > key(c); });
    }
    out.clear();
    out.reserve(total);
    using E = std::pair<uint64_t, uint32_t>;   // key, job
    auto lower = [](const E& a, const E& c) {   // the heap's top: largest key, lowest job
        return a.first < c.first || (a.first == c.first && a.second > c.second);
    };
    std::vector<E> heap;
    std::vector<size_t> at(lists.size(), 0);
    for (uint32_t j = 0; j < (uint32_t)lists.size(); ++j)
        if (!lists[j].empty()) heap.emplace_back(key(lists[j][0]), j);
    std::make_heap(heap.begin(), heap.end(), lower);
    while (!heap.empty()) {
        std::pop_heap(heap.begin(), heap.end(), lower);
        const uint32_t j = heap.back().second;
        heap.pop_back();
        out.push_back(lists[j][at[j]++]);
        if (at[j] < lists[j].size()) {
            heap.emplace_back(key(lists[j][at[j]]), j);
            std::push_heap(heap.begin(), heap.end(), lower);
        }
    }
}

void merge_jobs(const std::vector<const JobState*>& jss, const std::vector<const hpdex_matrix*>& ms,
                MergedJobs& out) {
    std::vector<std::vector<CountItem>> ci(jss.size());
    std::vector<std::vector<WorkItem>> gi(jss.size());
    out.jobs.clear();
    out.lds = 0;
    out.G = 2;
    for (size_t i = 0; i < jss.size(); ++i) {
        const JobState& js = *jss[i];
        const int32_t job = (int32_t)i;
        out.jobs.push_back(js.J);
        out.lds = std::max(out.lds, js.count_lds);
        out.G = std::max(out.G, (int)js.G);
        ci[i] = js.h_count_items;
        for (CountItem& it : ci[i]) it.flags |= (uint32_t)job << kCountJobShift;
        gi[i] = js.h_gen_items;
        for (WorkItem& w : gi[i]) w.job = job;
    }
    // largest first: count items by the entries they read, sort-path columns
    // by their length (the orders a size-ordered plan lays its lists out in)
    merge_ordered(ci, [](const CountItem& it) { return (uint64_t)it.n; }, out.count);
    merge_ordered(gi, [&](const WorkItem& w) {
        const int64_t* ip = ms[(size_t)w.job]->h_indptr.data();
        return (uint64_t)(ip[w.col + 1] - ip[w.col]);
    }, out.gen);
}

int upload_merged(const MergedJobs& mj, bool f64, const Tuning& tu, DevBuf& d_jobs,
                  DevBuf& d_count, DevBuf& d_gen, hipStream_t s, BatchLaunch& L, char* err,
                  size_t err_len) {
    auto put = [&](DevBuf& d, const void* src, size_t bytes) -> hipError_t {
        if (!(d.p && d.n >= bytes)) {
            const hipError_t e = d.alloc(bytes ? bytes : 16);
            if (e != hipSuccess) return e;
        }
        return bytes ? hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, s) : hipSuccess;
    };
    HIPCHK(put(d_jobs, mj.jobs.data(), mj.jobs.size() * sizeof(JobDesc)));
    HIPCHK(put(d_count, mj.count.data(), mj.count.size() * sizeof(CountItem)));
    HIPCHK(put(d_gen, mj.gen.data(), mj.gen.size() * sizeof(WorkItem)));
    HIPCHK(hipStreamSynchronize(s));   // the lists are pageable host memory
    L = BatchLaunch{};
    L.jobs = d_jobs.as<JobDesc>();
    L.count_items = d_count.as<CountItem>();
    L.n_count = (int)mj.count.size();
    L.gen_items = d_gen.as<WorkItem>();
    L.n_gen = (int)mj.gen.size();
    L.G = mj.G;
    L.f64 = f64;
    L.count_lds = mj.lds ? mj.lds : (uint32_t)tu.count_lds;
    L.count_block = tu.count_block;
    L.split_lds_max = g_count_ready ? (uint32_t)kSplitLdsMax : (uint32_t)kLdsBudget;
    return 0;
}

// The batch's launch sequence: per value type, the persistent count kernel
// over every plan's items and the sort-path kernels over every plan's other
// columns; then each exact plan's p-value lookup.
static int batch_enqueue(hpdex_batch* b, char* err, size_t err_len) {
    const hipStream_t s = b->st.s;
    for (int v = 0; v < 2; ++v)
        if (b->L[v].n_count > 0 || b->L[v].n_gen > 0) HIPCHK(launch_mwu_batch(b->L[v], s));
    for (hpdex_plan* p : b->plans)
        if (p->js.opts.method == 1 && exact_launch(p->js, s, false, err, err_len)) return 1;
    return 0;
}

}  // namespace hpdex_host

extern "C" {

int hpdex_batch_create(hpdex_plan* const* plans, size_t n_plans, hpdex_batch** out, char* err,
                       size_t err_len) {
    if (!out || (n_plans && !plans)) return set_err(err, err_len, "null argument");
    *out = nullptr;
    if (n_plans == 0) return set_err(err, err_len, "a batch needs at least one plan");
    if (n_plans >= ((size_t)1 << (32 - kCountJobShift)))
        return set_err(err, err_len, "too many plans in one batch");
    if (ensure_init(err, err_len)) return 1;
    std::unique_ptr<hpdex_batch> b(new hpdex_batch());
    b->plans.assign(plans, plans + n_plans);
    for (size_t i = 0; i < n_plans; ++i) {
        if (!plans[i]) return set_err(err, err_len, "null plan in batch");
        for (size_t k = 0; k < i; ++k)
            if (plans[k] == plans[i]) return set_err(err, err_len, "a plan appears twice in the batch");
        if (plans[i]->js.opts.method == 1 && plans[i]->js.opts.zero_handling == 0)
            return set_err(err, err_len,
                           "exact p with zero_handling none sizes its tables from the data: run "
                           "that plan on its own");
    }
    if (b->st.create() != hipSuccess) return set_err(err, err_len, "stream creation failed");
    const Tuning tu = tuning();
    for (int v = 0; v < 2; ++v) {
        std::vector<const JobState*> jss;
        std::vector<const hpdex_matrix*> ms;
        for (size_t i = 0; i < n_plans; ++i)
            if ((int)plans[i]->m->f64 == v) {
                jss.push_back(&plans[i]->js);
                ms.push_back(plans[i]->m);
            }
        if (jss.empty()) continue;
        MergedJobs mj;
        merge_jobs(jss, ms, mj);
        if (upload_merged(mj, v == 1, tu, b->d_jobs[v], b->d_count[v], b->d_gen[v], b->st.s,
                          b->L[v], err, err_len))
            return 1;
        b->n_count_items += b->L[v].n_count;
        b->n_gen_items += b->L[v].n_gen;
    }
    // one graph for the whole sequence, replayed by hpdex_batch_run
    if (hipStreamBeginCapture(b->st.s, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return set_err(err, err_len, "graph capture failed to start");
    const int rc = batch_enqueue(b.get(), err, err_len);
    const hipError_t e = hipStreamEndCapture(b->st.s, &b->graph);
    if (rc) return 1;
    if (e != hipSuccess) return set_err(err, err_len, hipGetErrorString(e));
    if (hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess)
        return set_err(err, err_len, "graph instantiation failed");
    *out = b.release();
    return 0;
}

int hpdex_batch_run(hpdex_batch* b, int iters, double* ms_total, char* err, size_t err_len) {
    if (!b) return set_err(err, err_len, "null batch");
    if (iters < 1) iters = 1;
    const hipStream_t s = b->st.s;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int rc = 0;
    (void)hipEventRecord(e0, s);
    for (int it = 0; it < iters && !rc; ++it) {
        const hipError_t e = hipGraphLaunch(b->exec, s);
        if (e != hipSuccess) rc = set_err(err, err_len, hipGetErrorString(e));
    }
    (void)hipEventRecord(e1, s);
    const hipError_t se = hipStreamSynchronize(s);
    if (!rc && se != hipSuccess) rc = set_err(err, err_len, hipGetErrorString(se));
    if (!rc && ms_total) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        *ms_total = ms;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    // each plan's matrix keeps its sorted reference after the first run
    for (hpdex_plan* p : b->plans)
        if (!rc) rc = refcache_fill(p->js, *p->m, p->st.s, err, err_len);
    return rc;
}

int hpdex_batch_get_info(hpdex_batch* b, int64_t* count_items, int64_t* split_columns) {
    if (!b) return 1;
    if (count_items) *count_items = b->n_count_items;
    if (split_columns) *split_columns = b->n_gen_items;
    return 0;
}

void hpdex_batch_destroy(hpdex_batch* b) { delete b; }

}  // extern "C"
