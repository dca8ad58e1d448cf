// hpdex_de.hip -- differential-expression post-processing on the device: fold
// change / log2 fold change from group means (src/hpdex/de.py:126-142) and the
// Benjamini-Hochberg FDR of every p-value (de.py:172-184, which calls
// scipy.stats.false_discovery_control(P, method="bh")).
//
// BH as scipy computes it: order the m p-values ascending, multiply the k-th
// (1-based) by fl(m / k) in FP64, take the running minimum from the largest
// down, put the values back in input order and clip to [0, 1].  Equal
// p-values end with equal adjusted values whatever their order inside the
// sort (the running minimum reaches back over the whole tie), so the sort
// need not be stable.  The sort is a device radix sort of the p-values' IEEE
// bits (non-negative doubles order as their bit patterns), paired with their
// positions (rocPRIM); the minimum runs as a reversed inclusive scan.
#include <cstring>

#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/reverse_iterator.hpp>

#include "hpdex_device.h"
#include "hpdex_internal.h"

namespace hpdex {

// fold_change[t*C + c] = mean[c*G + t+1] / mean[c*G], with the clip rules of
// de.py:136-141; log2 as numpy's (the FP64 log2 of the device library).
__global__ __launch_bounds__(256) void fold_change_kernel(const double* __restrict__ means,
                                                          int64_t C, int32_t T, double clip,
                                                          double* __restrict__ fc,
                                                          double* __restrict__ log2fc) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C * (int64_t)T) return;
    const int64_t t = i / C, c = i - t * C;
    const int64_t G = (int64_t)T + 1;
    const double ref = means[c * G], tar = means[c * G + t + 1];
    double f = tar / ref;
    if (ref < 1e-10) f = clip;
    if (tar < 1e-10) f = 1.0 / clip;
    fc[i] = f;
    log2fc[i] = log2(f);
}

// keys: IEEE bits of p (p in [0, 1]); a value outside [0, 1] (or NaN) raises
// the flag, as scipy raises for it.
__global__ __launch_bounds__(256) void fdr_keys_kernel(const double* __restrict__ p, int64_t n,
                                                       unsigned long long* __restrict__ keys,
                                                       uint32_t* __restrict__ idx,
                                                       uint32_t* __restrict__ bad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = p[i];
    if (!(v >= 0.0 && v <= 1.0)) {
        atomicOr(bad, 1u);
        keys[i] = 0ull;
    } else {
        keys[i] = (unsigned long long)__double_as_longlong(v);
    }
    idx[i] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void fdr_scale_kernel(const unsigned long long* __restrict__ keys,
                                                        int64_t n, double* __restrict__ adj) {
#pragma clang fp contract(off)
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double pk = __longlong_as_double((long long)keys[k]);
    const double q = (double)n / (double)(k + 1);
    adj[k] = pk * q;
}

__global__ __launch_bounds__(256) void fdr_scatter_kernel(const double* __restrict__ adjmin,
                                                          const uint32_t* __restrict__ idx,
                                                          int64_t n, const uint32_t* __restrict__ bad,
                                                          double* __restrict__ fdr) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double v = adjmin[k];
    v = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
    if (*bad) v = __longlong_as_double(0x7FF8000000000000ll);   // NaN: de.py's fallback
    fdr[idx[k]] = v;
}

// A fill kernel rather than a memset node, so the sequence can be captured.
__global__ void fdr_prepare_kernel(uint32_t* bad) {
    if (threadIdx.x == 0) *bad = 0u;
}

__global__ void copy_one_kernel(const double* in, double* out) {
    if (threadIdx.x == 0) *out = *in;
}

static inline unsigned grid1d(int64_t n) { return (unsigned)((n + 255) / 256); }

hipError_t launch_fold_change(const double* means, int64_t C, int32_t T, double clip, double* fc,
                              double* log2fc, hipStream_t s) {
    const int64_t n = C * (int64_t)T;
    if (n <= 0) return hipSuccess;
    fold_change_kernel<<<grid1d(n), 256, 0, s>>>(means, C, T, clip, fc, log2fc);
    return hipGetLastError();
}

size_t fdr_workspace_bytes(int64_t n) {
    if (n <= 1) return 0;
    size_t sort_tmp = 0, scan_tmp = 0;
    (void)rocprim::radix_sort_pairs(nullptr, sort_tmp, (unsigned long long*)nullptr,
                                    (unsigned long long*)nullptr, (uint32_t*)nullptr,
                                    (uint32_t*)nullptr, (size_t)n, 0, 64);
    (void)rocprim::inclusive_scan(nullptr, scan_tmp,
                                  rocprim::make_reverse_iterator((double*)nullptr + n),
                                  rocprim::make_reverse_iterator((double*)nullptr + n), (size_t)n,
                                  rocprim::minimum<double>());
    const size_t tmp = sort_tmp > scan_tmp ? sort_tmp : scan_tmp;
    auto a = [](size_t b) { return (b + 255) & ~(size_t)255; };
    // keys x2 (8 B), idx x2 (4 B), adj x2 (8 B), flag, rocPRIM scratch
    return 2 * a(8 * (size_t)n) + 2 * a(4 * (size_t)n) + 2 * a(8 * (size_t)n) + 256 + a(tmp);
}

// The same layout with rocPRIM's scratch bounded by 16 bytes per p-value
// plus 4 MiB instead of asked for (its size query needs the device): a radix
// sort of 8-byte keys with 4-byte values keeps at most one spare copy of both
// (12 bytes per value) and per-block digit counts; launch_fdr_bh refuses a
// workspace query above this bound.
size_t fdr_workspace_bound(int64_t n) {
    if (n <= 1) return 0;
    auto a = [](size_t b) { return (b + 255) & ~(size_t)255; };
    return 2 * a(8 * (size_t)n) + 2 * a(4 * (size_t)n) + 2 * a(8 * (size_t)n) + 256 +
           a(16 * (size_t)n + ((size_t)4 << 20));
}

hipError_t launch_fdr_bh(const double* p, int64_t n, double* fdr, void* ws, size_t ws_bytes,
                         hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (n == 1) {   // scipy returns a single p-value unchanged
        copy_one_kernel<<<1, 64, 0, s>>>(p, fdr);
        return hipGetLastError();
    }
    if (n > (int64_t)UINT32_MAX || ws_bytes < fdr_workspace_bytes(n)) return hipErrorInvalidValue;
    auto a = [](size_t b) { return (b + 255) & ~(size_t)255; };
    char* w = (char*)ws;
    auto* k0 = (unsigned long long*)w;  w += a(8 * (size_t)n);
    auto* k1 = (unsigned long long*)w;  w += a(8 * (size_t)n);
    auto* i0 = (uint32_t*)w;            w += a(4 * (size_t)n);
    auto* i1 = (uint32_t*)w;            w += a(4 * (size_t)n);
    auto* adj = (double*)w;             w += a(8 * (size_t)n);
    auto* amin = (double*)w;            w += a(8 * (size_t)n);
    auto* bad = (uint32_t*)w;           w += 256;
    const size_t used = (size_t)(w - (char*)ws);
    size_t tmp_bytes = ws_bytes - used;
    void* tmp = w;
    fdr_prepare_kernel<<<1, 64, 0, s>>>(bad);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    fdr_keys_kernel<<<grid1d(n), 256, 0, s>>>(p, n, k0, i0, bad);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    e = rocprim::radix_sort_pairs(tmp, tmp_bytes, k0, k1, i0, i1, (size_t)n, 0, 64, s);
    if (e != hipSuccess) return e;
    fdr_scale_kernel<<<grid1d(n), 256, 0, s>>>(k1, n, adj);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    tmp_bytes = ws_bytes - used;
    e = rocprim::inclusive_scan(tmp, tmp_bytes, rocprim::make_reverse_iterator(adj + n),
                                rocprim::make_reverse_iterator(amin + n), (size_t)n,
                                rocprim::minimum<double>(), s);
    if (e != hipSuccess) return e;
    fdr_scatter_kernel<<<grid1d(n), 256, 0, s>>>(amin, i1, n, bad, fdr);
    return hipGetLastError();
}

}  // namespace hpdex
