// hpdex_ingest.hip -- a resident CSC matrix built on the device from the
// layouts callers hold: CSR (AnnData's X, which src/hpdex/de.py:103-109
// converts with scipy's tocsc on the host), a dense row-major array (the same
// lines, csc_matrix(X)), and CSC with 32-bit indices (the wrapper's int64
// copy, src/hpdex/backend/mannwhitneyu_.py:59-62, is not needed on the device).
//
// CSR -> CSC: every stored entry gets its row from the row pointers, a stable
// radix sort of (column, entry) pairs orders entries by column and keeps the
// row order inside a column (what tocsc produces), the column pointers are
// lower bounds in the sorted column keys, and a gather writes rows and values
// (converted as the reference converts, (double)v or float) in CSC order.
//
// dense -> CSC: a tiled transpose through LDS gives column-major values, one
// workgroup per column counts its nonzeros, an exclusive scan gives the column
// pointers, and one workgroup per column writes its nonzeros in row order
// (ordered compaction: a ballot and a scan of wave counts per chunk).  Zeros,
// -0.0 included, are dropped and NaN kept, as csc_matrix does.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "hpdex_device.h"
#include "hpdex_internal.h"
#include "hpdex_mi355x.h"

namespace hpdex {

namespace {

inline unsigned grid_of(size_t n, unsigned block) {
    return (unsigned)((n + block - 1) / block);
}

// 32-bit row indices outside [0, R) become -1 (an ignored entry), as the
// 64-bit upload does
__global__ void rows_i32_check_kernel(int32_t* rows, size_t n, int32_t R) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t r = rows[i];
    if (r < 0 || r >= R) rows[i] = -1;
}

// A CSC matrix uploaded compactly: one 32-bit word per stored entry, the
// row's low 16 bits above the value as a 16-bit unsigned integer (float32
// values that are whole numbers in [0, 65535]), in place of the rows array.
// Rows of a column ascend with gaps below 2^16, so their high bits start at
// hi[c] (the column's first row's) and go up by one wherever the low bits
// fall.  One workgroup per column decodes it in place: chunks of 1,024
// entries, 4 per thread, the low bits through LDS (each entry's predecessor),
// a scan of the wrap flags, then rows and values written.
constexpr uint32_t kDecodeChunk = 1024u;
__global__ __launch_bounds__(256) void decode_compact_kernel(uint32_t* words, float* vals,
                                                             const int64_t* indptr,
                                                             const int32_t* hi) {
    __shared__ uint32_t los[kDecodeChunk];
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t carry_hi, carry_lo;
    const int64_t c = blockIdx.x;
    const int64_t p0 = indptr[c], p1 = indptr[c + 1];
    if (p0 >= p1) return;   // the same in every thread
    const uint32_t tid = threadIdx.x, lane = (uint32_t)lane_id(), w = tid / kWave;
    if (tid == 0) {
        carry_hi = (uint32_t)hi[c];
        carry_lo = 0xFFFFFFFFu;   // the column's first entry has no predecessor
    }
    for (int64_t base = p0; base < p1; base += kDecodeChunk) {
        __syncthreads();   // the carry of the chunk before; los free
        uint32_t wd[4];
        bool in[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int64_t p = base + tid * 4 + k;
            in[k] = p < p1;
            wd[k] = in[k] ? words[p] : 0u;
            los[tid * 4 + k] = wd[k] >> 16;
        }
        const uint32_t hi0 = carry_hi, prev0 = carry_lo;
        __syncthreads();
        uint32_t f[4], cnt = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t i = tid * 4 + k;
            const uint32_t prev = i ? los[i - 1] : prev0;
            f[k] = (in[k] && prev != 0xFFFFFFFFu && (wd[k] >> 16) < prev) ? 1u : 0u;
            cnt += f[k];
        }
        const uint32_t incl = wave_incl_scan_u32(cnt);
        if (lane == (uint32_t)kWave - 1u) wsum[w] = incl;
        __syncthreads();
        uint32_t woff = 0, total = 0;
#pragma unroll
        for (uint32_t x = 0; x < 4u; ++x) {
            woff += x < w ? wsum[x] : 0u;
            total += wsum[x];
        }
        uint32_t run = hi0 + woff + incl - cnt;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            run += f[k];
            if (!in[k]) continue;
            const int64_t p = base + tid * 4 + k;
            words[p] = (run << 16) | (wd[k] >> 16);
            vals[p] = (float)(wd[k] & 0xFFFFu);
        }
        __syncthreads();   // every thread has read the carry and los
        if (tid == 0) {
            const int64_t last = (p1 - base < (int64_t)kDecodeChunk ? p1 - base : kDecodeChunk) - 1;
            carry_hi = hi0 + total;
            carry_lo = los[last];
        }
    }
}

// row of every CSR entry: one wave per row, lanes over its entries
template <typename I>
__global__ __launch_bounds__(256) void csr_rows_kernel(const I* indptr, size_t R, int32_t* row_of) {
    const size_t r = (size_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (r >= R) return;
    const int64_t e = (int64_t)indptr[r + 1];
    for (int64_t p = (int64_t)indptr[r] + lane_id(); p < e; p += kWave) row_of[p] = (int32_t)r;
}

// column keys and entry positions for the sort; a column index outside
// [0, C) raises the flag (the host reports it)
template <typename I>
__global__ void csr_keys_kernel(const I* cols, size_t nnz, int64_t C, uint32_t* keys,
                                uint32_t* pos, uint32_t* bad) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnz) return;
    const int64_t c = (int64_t)cols[i];
    if (c < 0 || c >= C) {
        atomicOr(bad, 1u);
        keys[i] = 0u;
    } else {
        keys[i] = (uint32_t)c;
    }
    pos[i] = (uint32_t)i;
}

// indptr[c] = first position of a key >= c in the sorted keys (c = C: nnz)
__global__ void csc_indptr_kernel(const uint32_t* keys, size_t nnz, int64_t C, int64_t* indptr) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c > C) return;
    size_t lo = 0, hi = nnz;
    while (lo < hi) {
        const size_t mid = lo + (hi - lo) / 2;
        if ((int64_t)keys[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    indptr[c] = (int64_t)lo;
}

template <typename Tin, typename Tout>
__global__ void csr_gather_kernel(const Tin* vals, const int32_t* row_of, const uint32_t* pos,
                                  size_t nnz, Tout* vals_out, int32_t* rows_out) {
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const uint32_t p = pos[k];
    vals_out[k] = (Tout)vals[p];
    rows_out[k] = row_of[p];
}

// dense row-major R x C -> column-major, converted; 64 x 64 tiles, one per
// workgroup (tile index = blockIdx.x, column tiles fastest)
constexpr int kTile = 64;
template <typename Tin, typename Tout>
__global__ __launch_bounds__(256) void transpose_kernel(const Tin* in, size_t R, size_t C,
                                                        size_t col_tiles, Tout* out) {
    __shared__ Tout tile[kTile][kTile + 1];
    const size_t t = blockIdx.x;
    const size_t c0 = (t % col_tiles) * kTile, r0 = (t / col_tiles) * kTile;
    const int tx = (int)(threadIdx.x % kTile), ty = (int)(threadIdx.x / kTile);   // ty in [0, 4)
#pragma unroll
    for (int k = 0; k < kTile / 4; ++k) {
        const size_t r = r0 + (size_t)(ty + 4 * k), c = c0 + (size_t)tx;
        if (r < R && c < C) tile[ty + 4 * k][tx] = (Tout)in[r * C + c];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kTile / 4; ++k) {
        const size_t c = c0 + (size_t)(ty + 4 * k), r = r0 + (size_t)tx;
        if (r < R && c < C) out[c * R + r] = tile[tx][ty + 4 * k];
    }
}

// nonzeros per column of the column-major copy: one workgroup per column
template <typename T>
__global__ __launch_bounds__(256) void count_nonzero_kernel(const T* colmajor, size_t R,
                                                            int64_t* cnt) {
    __shared__ uint64_t red[16];
    const size_t c = blockIdx.x;
    const T* col = colmajor + c * R;
    uint64_t k = 0;
    for (size_t r = threadIdx.x; r < R; r += blockDim.x) k += (col[r] != T(0)) ? 1u : 0u;
    k = block_sum_u64(k, red);
    if (threadIdx.x == 0) cnt[c] = (int64_t)k;
}

// nonzeros of each column in row order, at indptr[c]: chunks of blockDim
// rows, a ballot per wave and a scan of the wave counts place each one
template <typename T>
__global__ __launch_bounds__(256) void compact_kernel(const T* colmajor, size_t R,
                                                      const int64_t* indptr, T* vals_out,
                                                      int32_t* rows_out) {
    __shared__ uint32_t wcnt[16];
    __shared__ uint32_t woff[17];
    const size_t c = blockIdx.x;
    const T* col = colmajor + c * R;
    const int tid = (int)threadIdx.x, nw = (int)(blockDim.x / kWave), w = tid / kWave;
    int64_t out = indptr[c];
    for (size_t base = 0; base < R; base += blockDim.x) {
        const size_t r = base + (size_t)tid;
        const T v = r < R ? col[r] : T(0);
        const bool nz = v != T(0);
        const uint64_t bal = __ballot(nz ? 1 : 0);
        const uint32_t rank = (uint32_t)__builtin_popcountll(bal & ((1ull << lane_id()) - 1ull));
        if (lane_id() == 0) wcnt[w] = (uint32_t)__builtin_popcountll(bal);
        __syncthreads();
        if (tid == 0) {
            uint32_t s = 0;
            for (int i = 0; i < nw; ++i) { woff[i] = s; s += wcnt[i]; }
            woff[16] = s;
        }
        __syncthreads();
        if (nz) {
            const int64_t at = out + woff[w] + rank;
            vals_out[at] = v;
            rows_out[at] = (int32_t)r;
        }
        out += woff[16];
        __syncthreads();
    }
}

template <typename Tin>
hipError_t gather_dispatch(bool f64, const void* vals, const int32_t* row_of, const uint32_t* pos,
                           size_t nnz, void* vals_out, int32_t* rows_out, hipStream_t s) {
    if (f64)
        csr_gather_kernel<Tin, double><<<grid_of(nnz, 256), 256, 0, s>>>(
            (const Tin*)vals, row_of, pos, nnz, (double*)vals_out, rows_out);
    else
        csr_gather_kernel<Tin, float><<<grid_of(nnz, 256), 256, 0, s>>>(
            (const Tin*)vals, row_of, pos, nnz, (float*)vals_out, rows_out);
    return hipGetLastError();
}

template <typename Tin>
hipError_t transpose_dispatch(bool f64, const void* in, size_t R, size_t C, void* out,
                              hipStream_t s) {
    const size_t ct = (C + kTile - 1) / kTile, rt = (R + kTile - 1) / kTile;
    if (ct * rt > 0x7FFFFFFFull) return hipErrorInvalidValue;
    if (f64)
        transpose_kernel<Tin, double><<<(unsigned)(ct * rt), 256, 0, s>>>((const Tin*)in, R, C, ct,
                                                                          (double*)out);
    else
        transpose_kernel<Tin, float><<<(unsigned)(ct * rt), 256, 0, s>>>((const Tin*)in, R, C, ct,
                                                                         (float*)out);
    return hipGetLastError();
}

// the reference's dtype codes (macro.hpp:4-14, module.cpp:144-166)
template <template <typename> class F, typename... A>
hipError_t by_dtype(int dtype, A... a) {
    switch (dtype) {
    case HPDEX_INT8: return F<int8_t>::run(a...);
    case HPDEX_UINT8: return F<uint8_t>::run(a...);
    case HPDEX_INT16: return F<int16_t>::run(a...);
    case HPDEX_UINT16: return F<uint16_t>::run(a...);
    case HPDEX_INT32: return F<int32_t>::run(a...);
    case HPDEX_UINT32: return F<uint32_t>::run(a...);
    case HPDEX_INT64: return F<int64_t>::run(a...);
    case HPDEX_UINT64: return F<uint64_t>::run(a...);
    case HPDEX_FLOAT32: return F<float>::run(a...);
    case HPDEX_FLOAT64: return F<double>::run(a...);
    default: return hipErrorInvalidValue;
    }
}
template <typename T> struct Gather {
    static hipError_t run(bool f64, const void* vals, const int32_t* row_of, const uint32_t* pos,
                          size_t nnz, void* vals_out, int32_t* rows_out, hipStream_t s) {
        return gather_dispatch<T>(f64, vals, row_of, pos, nnz, vals_out, rows_out, s);
    }
};
template <typename T> struct Transpose {
    static hipError_t run(bool f64, const void* in, size_t R, size_t C, void* out, hipStream_t s) {
        return transpose_dispatch<T>(f64, in, R, C, out, s);
    }
};

size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

hipError_t launch_rows_i32_check(int32_t* rows, size_t n, int64_t R, hipStream_t s) {
    if (n == 0) return hipSuccess;
    rows_i32_check_kernel<<<grid_of(n, 256), 256, 0, s>>>(rows, n, (int32_t)R);
    return hipGetLastError();
}

hipError_t launch_decode_compact(uint32_t* words, float* vals, const int64_t* indptr,
                                 const int32_t* hi, size_t C, hipStream_t s) {
    if (C == 0) return hipSuccess;
    decode_compact_kernel<<<(unsigned)C, 256, 0, s>>>(words, vals, indptr, hi);
    return hipGetLastError();
}

size_t csr_workspace_bytes(size_t nnz) {
    size_t sort_tmp = 0;
    (void)rocprim::radix_sort_pairs(nullptr, sort_tmp, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                    (uint32_t*)nullptr, (uint32_t*)nullptr, nnz, 0, 32);
    // keys x2, positions x2, row of each entry, flag, sort scratch
    return 5 * up256(4 * nnz) + 256 + up256(sort_tmp);
}

hipError_t launch_csr_to_csc(int dtype, bool f64, const void* vals, const void* cols,
                             const void* indptr, int index_bits, size_t R, size_t C, size_t nnz,
                             void* vals_out, int32_t* rows_out, int64_t* indptr_out,
                             uint32_t* bad_host, void* ws, size_t ws_bytes, hipStream_t s) {
    *bad_host = 0u;
    if (nnz > (size_t)UINT32_MAX || ws_bytes < csr_workspace_bytes(nnz)) return hipErrorInvalidValue;
    char* w = (char*)ws;
    auto* k0 = (uint32_t*)w;     w += up256(4 * nnz);
    auto* k1 = (uint32_t*)w;     w += up256(4 * nnz);
    auto* p0 = (uint32_t*)w;     w += up256(4 * nnz);
    auto* p1 = (uint32_t*)w;     w += up256(4 * nnz);
    auto* row_of = (int32_t*)w;  w += up256(4 * nnz);
    auto* bad = (uint32_t*)w;    w += 256;
    size_t tmp_bytes = ws_bytes - (size_t)(w - (char*)ws);
    hipError_t e = hipMemsetAsync(bad, 0, 4, s);
    if (e != hipSuccess) return e;
    if (nnz) {
        const unsigned rg = grid_of(R, 256 / kWave);
        if (index_bits == 32) {
            csr_rows_kernel<int32_t><<<rg, 256, 0, s>>>((const int32_t*)indptr, R, row_of);
            csr_keys_kernel<int32_t><<<grid_of(nnz, 256), 256, 0, s>>>((const int32_t*)cols, nnz,
                                                                       (int64_t)C, k0, p0, bad);
        } else {
            csr_rows_kernel<int64_t><<<rg, 256, 0, s>>>((const int64_t*)indptr, R, row_of);
            csr_keys_kernel<int64_t><<<grid_of(nnz, 256), 256, 0, s>>>((const int64_t*)cols, nnz,
                                                                       (int64_t)C, k0, p0, bad);
        }
        if ((e = hipGetLastError()) != hipSuccess) return e;
        int end_bit = 1;
        while (end_bit < 32 && ((uint64_t)1 << end_bit) < (uint64_t)C) ++end_bit;
        e = rocprim::radix_sort_pairs(w, tmp_bytes, k0, k1, p0, p1, nnz, 0, end_bit, s);
        if (e != hipSuccess) return e;
        e = by_dtype<Gather>(dtype, f64, vals, (const int32_t*)row_of, (const uint32_t*)p1, nnz,
                             vals_out, rows_out, s);
        if (e != hipSuccess) return e;
    }
    csc_indptr_kernel<<<grid_of(C + 1, 256), 256, 0, s>>>(k1, nnz, (int64_t)C, indptr_out);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    e = hipMemcpyAsync(bad_host, bad, 4, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

size_t dense_workspace_bytes(size_t R, size_t C, bool f64) {
    size_t scan_tmp = 0;
    (void)rocprim::exclusive_scan(nullptr, scan_tmp, (int64_t*)nullptr, (int64_t*)nullptr,
                                  (int64_t)0, C + 1, rocprim::plus<int64_t>());
    return up256(R * C * (f64 ? 8 : 4)) + up256(8 * (C + 1)) + up256(scan_tmp);
}

hipError_t launch_dense_count(int dtype, bool f64, const void* dense, size_t R, size_t C,
                              int64_t* indptr_out, void* ws, size_t ws_bytes, hipStream_t s) {
    if (ws_bytes < dense_workspace_bytes(R, C, f64)) return hipErrorInvalidValue;
    char* w = (char*)ws;
    void* colmajor = w;             w += up256(R * C * (f64 ? 8 : 4));
    auto* cnt = (int64_t*)w;        w += up256(8 * (C + 1));
    size_t tmp_bytes = ws_bytes - (size_t)(w - (char*)ws);
    hipError_t e = hipSuccess;
    if (R && C) {
        e = by_dtype<Transpose>(dtype, f64, dense, R, C, colmajor, s);
        if (e != hipSuccess) return e;
        if (f64) count_nonzero_kernel<double><<<(unsigned)C, 256, 0, s>>>((const double*)colmajor, R, cnt);
        else count_nonzero_kernel<float><<<(unsigned)C, 256, 0, s>>>((const float*)colmajor, R, cnt);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if ((e = hipMemsetAsync(cnt + C, 0, 8, s)) != hipSuccess) return e;
    if (C == 0) return hipMemsetAsync(indptr_out, 0, 8, s);
    return rocprim::exclusive_scan(w, tmp_bytes, cnt, indptr_out, (int64_t)0, C + 1,
                                   rocprim::plus<int64_t>(), s);
}

hipError_t launch_dense_compact(bool f64, size_t R, size_t C, const int64_t* indptr, void* vals_out,
                                int32_t* rows_out, void* ws, hipStream_t s) {
    if (R == 0 || C == 0) return hipSuccess;
    if (f64)
        compact_kernel<double><<<(unsigned)C, 256, 0, s>>>((const double*)ws, R, indptr,
                                                           (double*)vals_out, rows_out);
    else
        compact_kernel<float><<<(unsigned)C, 256, 0, s>>>((const float*)ws, R, indptr,
                                                          (float*)vals_out, rows_out);
    return hipGetLastError();
}

}  // namespace hpdex
