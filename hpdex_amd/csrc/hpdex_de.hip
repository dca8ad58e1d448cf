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

// ---------------------------------------------------------------------------
// The numba pipeline (pden): src/hpdex/backend/de_.py with the kernels of
// src/hpdex/backend/pykernel_.py, from the Mann-Whitney pass's reference U and
// raw tie sums under zero_handling "mix" (the pipeline densifies the matrix,
// so zeros rank at their true position and n is the group's row count).
// ---------------------------------------------------------------------------

// fold change with de_.py:189-247 on numpy's means (de_.py:640-652): a float32
// matrix's mean is the float32 sum divided in float64 by the row count
// (numpy's true_divide by an intp count) and rounded to float32, and the
// ratio, the clip values and log2 are float32 (log2 as the float64 log2
// rounded to float32); any other dtype's mean and ratio are float64.  The target-zero rule runs first, so the reference-zero
// rule wins when both hold.
__global__ __launch_bounds__(256) void pden_fold_change_kernel(
    const double* __restrict__ sums, const uint32_t* __restrict__ group_rows, int64_t C, int32_t T,
    double clip, int mean_f32, double* __restrict__ fc, double* __restrict__ log2fc) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C * (int64_t)T) return;
    const int64_t t = i / C, c = i - t * C;
    const int64_t G = (int64_t)T + 1;
    const double nr = (double)group_rows[0], nt = (double)group_rows[t + 1];
    if (mean_f32) {
        const float mr = (float)((double)(float)sums[c * G] / nr);
        const float mt = (float)((double)(float)sums[c * G + t + 1] / nt);
        float f = mt / mr;
        if ((double)mt < 1e-10) f = (float)(1.0 / clip);
        if ((double)mr < 1e-10) f = (float)clip;
        fc[i] = (double)f;
        log2fc[i] = (double)(float)log2((double)f);
    } else {
        const double mr = sums[c * G] / nr, mt = sums[c * G + t + 1] / nt;
        double f = mt / mr;
        if (mt < 1e-10) f = 1.0 / clip;
        if (mr < 1e-10) f = clip;
        fc[i] = f;
        log2fc[i] = log2(f);
    }
}

// Sums of every (column, group) in the order numpy's axis-0 mean adds them
// (de_.py:640-643 on the dense matrix: row by row, in the accumulator type,
// float32 for a float32 matrix and float64 for any other): one wave per
// column walks its entries in storage order (row order for a canonical CSC
// and for every matrix built on the device) 64 at a time; inside a batch the
// lanes of one group add in lane order, one lane per group and round.  The
// implicit zeros change no sum.  Rows outside [0, R) or any group are skipped.
template <typename V, typename A>
__global__ __launch_bounds__(64) void pden_group_sum_kernel(
    const V* __restrict__ vals, const int32_t* __restrict__ rows, const int64_t* __restrict__ indptr,
    const int32_t* __restrict__ gid, int32_t R, int32_t G, int32_t S, double* __restrict__ out) {
#pragma clang fp contract(off)
    extern __shared__ __align__(16) unsigned char smem[];
    // workgroup (c, y): column c, groups [y*S, y*S + S) (each slice of groups
    // reads the column once)
    const int32_t g0 = (int32_t)blockIdx.y * S;
    const int32_t gs = min(S, G - g0);
    A* sum = (A*)smem;
    int32_t* owner = (int32_t*)(sum + gs);
    const int64_t c = blockIdx.x;
    const int lane = (int)threadIdx.x;
    for (int g = lane; g < gs; g += 64) { sum[g] = A(0); owner[g] = 64; }
    __syncthreads();
    const int64_t p0 = indptr[c], p1 = indptr[c + 1];
    for (int64_t pb = p0; pb < p1; pb += 64) {
        const int64_t p = pb + lane;
        int32_t g = -1;
        A v = A(0);
        if (p < p1) {
            const int32_t r = rows[p];
            if (r >= 0 && r < R) {
                const int32_t gg = gid[r] - g0;   // the slice's own groups only
                if (gg >= 0 && gg < gs) { g = gg; v = (A)vals[p]; }
            }
        }
        bool pending = g >= 0;
        while (__ballot(pending) != 0ull) {
            if (pending) atomicMin(&owner[g], lane);
            __syncthreads();
            const bool win = pending && owner[g] == lane;
            if (win) sum[g] = sum[g] + v;
            __syncthreads();
            if (win) { owner[g] = 64; pending = false; }
            __syncthreads();
        }
    }
    for (int g = lane; g < gs; g += 64) out[c * G + g0 + g] = (double)sum[g];
}

// Order-preserving 64-bit key of a double, for atomic min / max.
__device__ inline unsigned long long order_key(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ inline double key_value(unsigned long long k) {
    const unsigned long long b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
    return __longlong_as_double((long long)b);
}

__global__ void pden_clear_kernel(unsigned long long* kmin, unsigned long long* kmax, int64_t nk,
                                  uint32_t* tflag, int64_t nf) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nk || i < nf;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (i < nk) { kmin[i] = ~0ull; kmax[i] = 0ull; }
        if (i < nf) tflag[i] = 0u;
    }
}

// Smallest and largest value of every group over each gene chunk, zeros
// included (a group with fewer stored entries in a column than rows holds a
// zero there), as the histogram kernel's value range (pykernel_.py:457-464)
// sees the dense chunk.  Workgroup (c, slice): column c, groups g0 .. g0+S-1
// in LDS.  Non-finite values are left out.
template <typename V>
__global__ __launch_bounds__(256) void pden_range_kernel(
    const V* __restrict__ vals, const int32_t* __restrict__ rows, const int64_t* __restrict__ indptr,
    const int32_t* __restrict__ gid, const uint32_t* __restrict__ group_rows, int32_t R, int32_t G,
    int32_t S, int64_t chunk, unsigned long long* kmin, unsigned long long* kmax) {
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned long long* mn = (unsigned long long*)smem;
    unsigned long long* mx = mn + S;
    uint32_t* cnt = (uint32_t*)(mx + S);
    const int64_t c = blockIdx.x;
    const int32_t g0 = (int32_t)blockIdx.y * S;
    const int32_t ns = (G - g0) < S ? (G - g0) : S;
    for (int x = threadIdx.x; x < ns; x += blockDim.x) { mn[x] = ~0ull; mx[x] = 0ull; cnt[x] = 0u; }
    __syncthreads();
    const int64_t p0 = indptr[c], p1 = indptr[c + 1];
    for (int64_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
        const int32_t r = rows[p];
        if (r < 0 || r >= R) continue;
        const int32_t g = gid[r] - g0;
        if (g < 0 || g >= ns) continue;
        const double v = (double)vals[p];
        atomicAdd(&cnt[g], 1u);
        if (!__builtin_isfinite(v)) continue;
        const unsigned long long k = order_key(v);
        atomicMin(&mn[g], k);
        atomicMax(&mx[g], k);
    }
    __syncthreads();
    const int64_t b = (c / chunk) * (int64_t)G + g0;
    const unsigned long long kz = order_key(0.0);
    for (int x = threadIdx.x; x < ns; x += blockDim.x) {
        unsigned long long lo = mn[x], hi = mx[x];
        if (cnt[x] < group_rows[g0 + x]) {
            lo = lo < kz ? lo : kz;
            hi = hi > kz ? hi : kz;
        }
        if (lo <= hi) {
            atomicMin(&kmin[b + x], lo);
            atomicMax(&kmax[b + x], hi);
        }
    }
}

// tflag[t*nchunk + k] = some gene of chunk k has a tie run (c > 1) for target
// t: the merges' has_ties (pykernel_.py:64-66, 379-382).
__global__ __launch_bounds__(256) void pden_ties_kernel(const double* __restrict__ tie, int64_t C,
                                                        int32_t T, int64_t chunk, int64_t nchunk,
                                                        uint32_t* tflag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C * (int64_t)T) return;
    if (!(tie[i] > 0.0)) return;
    const int64_t c = i / T, t = i - c * T;
    atomicOr(&tflag[t * nchunk + c / chunk], 1u);
}

// The method of each (target, chunk): use_asymptotic None -> asymptotic when
// the chunk has a tie or min(n_ref, n_tar) > 8 (pykernel_.py:267-271,
// 489-493); forced otherwise.  With the histogram kernel, a chunk whose value
// range (vmax - vmin + 1, pykernel_.py:457-464) passes max_bins falls back to
// the float merge with the asymptotic p forced (pykernel_.py:466-480).
__global__ __launch_bounds__(256) void pden_decide_kernel(
    const uint32_t* __restrict__ tflag, const unsigned long long* __restrict__ kmin,
    const unsigned long long* __restrict__ kmax, const uint32_t* __restrict__ group_rows, int32_t T,
    int64_t nchunk, int force, int use_hist, double max_bins, uint8_t* __restrict__ exact) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)T * nchunk) return;
    const int64_t t = i / nchunk, k = i - t * nchunk;
    const uint32_t n_ref = group_rows[0], n_tar = group_rows[t + 1];
    const int64_t G = (int64_t)T + 1;
    bool asym;
    if (force >= 0) asym = force == 1;
    else asym = tflag[i] != 0u || (n_ref < n_tar ? n_ref : n_tar) > 8u;
    if (use_hist) {
        double kp1 = 1.0;
        if (n_ref > 0u && n_tar > 0u) {
            const unsigned long long a0 = kmin[k * G], a1 = kmin[k * G + t + 1];
            const unsigned long long b0 = kmax[k * G], b1 = kmax[k * G + t + 1];
            const unsigned long long lo = a0 < a1 ? a0 : a1, hi = b0 > b1 ? b0 : b1;
            if (lo <= hi) kp1 = key_value(hi) - key_value(lo) + 1.0;
        }
        if (kp1 > max_bins) asym = true;
    }
    exact[i] = asym ? 0 : 1;
}

// Target U and p of every pair, target-major (the row order of de_.py:414-
// 428 and 487-502).  U_tar = n_ref * n_tar - U_ref, the merge's rank sum of
// the target minus n_tar (n_tar + 1) / 2 (pykernel_.py:74-76, 394-396).
// Asymptotic p in the order of operations of pykernel_.py:114-145 (n1 =
// n_tar, n2 = n_ref, its 1/sqrt(2) literal); exact p = 2 sf[rint(max(U_tar,
// U_ref))], clipped, with sf[k] = P(U >= k) from the table's counts
// (pykernel_.py:281-287, 501-507, 150-199; sf[k] = cdf[U - k] by the null's
// symmetry: the same sum in the same order).
__global__ __launch_bounds__(256) void pden_finish_kernel(PdenJob P) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.C * (int64_t)P.T) return;
    const int64_t t = i / P.C, c = i - t * P.C;
    const int64_t pair = c * P.T + t;
    const int64_t n2 = P.group_rows[0], n1 = P.group_rows[t + 1];
    const int64_t nm = n1 * n2;
    const double Uref = P.U1[pair];
    const double Utar = (double)nm - Uref;
    double p = 1.0;
    if (P.exact[t * P.nchunk + c / P.chunk]) {
        const ExactTable tb = P.tables[P.tidx[t]];
        const int64_t Umax = (int64_t)tb.n1 * tb.n2, W = Umax + 1;
        const double* base = P.tabbuf + tb.offset;
        const double hi = Utar > Uref ? Utar : Uref;
        int64_t u = (int64_t)rint(hi);
        u = u < 0 ? 0 : (u > Umax ? Umax : u);
        const double sf = base[1 + W + (Umax - u)] / base[0];
        p = 2.0 * sf;
        p = p < 0.0 ? 0.0 : (p > 1.0 ? 1.0 : p);
    } else {
        const int64_t N = n1 + n2;
        if (!(N <= 1 || n1 == 0 || n2 == 0)) {
            const double mu = 0.5 * (double)n1 * (double)n2;
            const double base = (double)nm * ((double)N + 1.0) / 12.0;
            const bool use_tie = P.tie_corr && N > 1;
            const double k_tie =
                use_tie ? (double)nm / (12.0 * (double)N * ((double)N - 1.0)) : 0.0;
            const double sigma2 = use_tie ? base - k_tie * P.tie[pair] : base;
            if (sigma2 <= 0.0 || !__builtin_isfinite(sigma2)) {
                p = 1.0;
            } else {
                double num = Utar - mu;
                if (P.cc != 0.0 && num != 0.0) num = num > 0.0 ? num - 0.5 : num + 0.5;
                const double zabs = fabs(num) / sqrt(sigma2);
                double pj = erfc(zabs * 0.7071067811865475);
                if (!__builtin_isfinite(pj) || pj > 1.0) pj = 1.0;
                else if (pj < 0.0) pj = 0.0;
                p = pj;
            }
        }
    }
    P.U_out[i] = Utar;
    P.P_out[i] = p;
}

hipError_t launch_pden_fold_change(const double* sums, const uint32_t* group_rows, int64_t C,
                                   int32_t T, double clip, bool mean_f32, double* fc,
                                   double* log2fc, hipStream_t s) {
    const int64_t n = C * (int64_t)T;
    if (n <= 0) return hipSuccess;
    pden_fold_change_kernel<<<grid1d(n), 256, 0, s>>>(sums, group_rows, C, T, clip,
                                                      mean_f32 ? 1 : 0, fc, log2fc);
    return hipGetLastError();
}

hipError_t launch_pden_group_sums(const void* vals, bool f64, bool acc_f32, const int32_t* rows,
                                  const int64_t* indptr, const int32_t* gid, int64_t C, int32_t R,
                                  int32_t G, double* sums, hipStream_t s) {
    if (C <= 0 || G <= 0) return hipSuccess;
    // a sum and an owner word per group: slices of as many groups as fit
    const size_t per = (acc_f32 ? 4 : 8) + 4;
    const int32_t S = (size_t)G * per <= (size_t)kLdsBudget ? G : (int32_t)(kLdsBudget / per);
    const dim3 grid((unsigned)C, (unsigned)((G + S - 1) / S));
    const size_t lds = (size_t)S * per;
    if (f64) {
        if (acc_f32) return hipErrorInvalidValue;   // a float32 matrix is stored as float
        pden_group_sum_kernel<double, double><<<grid, 64, lds, s>>>(
            (const double*)vals, rows, indptr, gid, R, G, S, sums);
    } else if (acc_f32) {
        pden_group_sum_kernel<float, float><<<grid, 64, lds, s>>>(
            (const float*)vals, rows, indptr, gid, R, G, S, sums);
    } else {
        pden_group_sum_kernel<float, double><<<grid, 64, lds, s>>>(
            (const float*)vals, rows, indptr, gid, R, G, S, sums);
    }
    return hipGetLastError();
}

hipError_t launch_pden_decide(const void* vals, bool f64, const int32_t* rows,
                              const int64_t* indptr, const int32_t* gid,
                              const uint32_t* group_rows, int64_t C, int32_t R, int32_t T,
                              const double* tie, int64_t chunk, int force, bool use_hist,
                              double max_bins, unsigned long long* kmin, unsigned long long* kmax,
                              uint32_t* tflag, uint8_t* exact, hipStream_t s) {
    if (C <= 0 || T <= 0 || chunk <= 0) return hipSuccess;
    const int64_t nchunk = (C + chunk - 1) / chunk;
    const int32_t G = T + 1;
    const int64_t nk = use_hist ? nchunk * (int64_t)G : 0, nf = (int64_t)T * nchunk;
    const int64_t nmax = nk > nf ? nk : nf;
    pden_clear_kernel<<<(unsigned)((nmax + 255) / 256 < 4096 ? (nmax + 255) / 256 : 4096), 256, 0,
                        s>>>(kmin, kmax, nk, tflag, nf);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    pden_ties_kernel<<<grid1d(C * (int64_t)T), 256, 0, s>>>(tie, C, T, chunk, nchunk, tflag);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (use_hist) {
        const int32_t S = (int32_t)(kLdsBudget / 20 < G ? kLdsBudget / 20 : G);
        const dim3 grid((unsigned)C, (unsigned)((G + S - 1) / S));
        const size_t lds = (size_t)S * 20;
        if (f64)
            pden_range_kernel<double><<<grid, 256, lds, s>>>((const double*)vals, rows, indptr, gid,
                                                             group_rows, R, G, S, chunk, kmin, kmax);
        else
            pden_range_kernel<float><<<grid, 256, lds, s>>>((const float*)vals, rows, indptr, gid,
                                                            group_rows, R, G, S, chunk, kmin, kmax);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    pden_decide_kernel<<<grid1d(nf), 256, 0, s>>>(tflag, kmin, kmax, group_rows, T, nchunk, force,
                                                  use_hist ? 1 : 0, max_bins, exact);
    return hipGetLastError();
}

hipError_t launch_pden_finish(const PdenJob& P, hipStream_t s) {
    const int64_t n = P.C * (int64_t)P.T;
    if (n <= 0) return hipSuccess;
    pden_finish_kernel<<<grid1d(n), 256, 0, s>>>(P);
    return hipGetLastError();
}

}  // namespace hpdex
