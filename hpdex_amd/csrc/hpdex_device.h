// hpdex_device.h -- device helpers for the gfx950 Mann-Whitney U kernels:
// order-preserving keys, 64-lane wave primitives, register bitonic sort,
// workgroup scans and the FP64 asymptotic p-value.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hpdex {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Order-preserving keys.  Values are compared as the reference compares them:
// promoted to double and ordered by '<' (src/cpp/mannwhitneyu.cpp:629-640,
// :920).  Zeros never become keys (they are counted per group, see the zero
// modes below), so +0.0/-0.0 need no special case, and non-finite values are
// dropped before keying (is_valid_value, :611-615).
// ---------------------------------------------------------------------------
template <typename V> struct KeyOf;

template <> struct KeyOf<float> {
    using K = uint32_t;
    static constexpr K kSentinel = 0xFFFFFFFFu;   // above every finite key
    static constexpr K kPosBit = 0x80000000u;      // keys >= this are positive
    __device__ static inline K key(float v) {
        uint32_t b = __float_as_uint(v);
        return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
    }
    // the value of a key (key's inverse)
    __device__ static inline float value(K k) {
        return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
    }
    __device__ static inline bool finite(float v) { return __builtin_isfinite(v); }
    // integral and small enough to bin exactly
    __device__ static inline bool small_int(float v, int& iv) {
        if (!(__builtin_fabsf(v) < 1073741824.f)) return false;
        iv = (int)v;
        return (float)iv == v;
    }
    // |v| <= lim, false for NaN
    __device__ static inline bool abs_le(float v, float lim) { return __builtin_fabsf(v) <= lim; }
};

template <> struct KeyOf<double> {
    using K = uint64_t;
    static constexpr K kSentinel = 0xFFFFFFFFFFFFFFFFull;
    static constexpr K kPosBit = 0x8000000000000000ull;
    __device__ static inline K key(double v) {
        uint64_t b = (uint64_t)__double_as_longlong(v);
        return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
    }
    __device__ static inline double value(K k) {
        return __longlong_as_double(
            (long long)((k & 0x8000000000000000ull) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k));
    }
    __device__ static inline bool finite(double v) { return __builtin_isfinite(v); }
    __device__ static inline bool small_int(double v, int& iv) {
        if (!(__builtin_fabs(v) < 1073741824.0)) return false;
        iv = (int)v;
        return (double)iv == v;
    }
    __device__ static inline bool abs_le(double v, double lim) { return __builtin_fabs(v) <= lim; }
};

// ---------------------------------------------------------------------------
// Wave primitives (wave64).
// ---------------------------------------------------------------------------
// Lane and thread indices pass through an empty volatile asm so that values
// derived from them are recomputed where used instead of being hoisted out of
// the persistent kernel's work loop (hoisting them pins ~80 extra VGPRs).
__device__ inline int lane_id() {
    int l = __lane_id();
    asm volatile("" : "+v"(l));
    return l;
}
__device__ inline int tid_x() {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

__device__ inline uint32_t shfl_xor_u32(uint32_t v, int m) {
    return (uint32_t)__shfl_xor((int)v, m, kWave);
}
__device__ inline uint64_t shfl_xor_u64(uint64_t v, int m) {
    uint32_t lo = shfl_xor_u32((uint32_t)v, m), hi = shfl_xor_u32((uint32_t)(v >> 32), m);
    return ((uint64_t)hi << 32) | lo;
}
__device__ inline uint32_t shfl_u32(uint32_t v, int src) { return (uint32_t)__shfl((int)v, src, kWave); }
__device__ inline uint64_t shfl_u64(uint64_t v, int src) {
    uint32_t lo = shfl_u32((uint32_t)v, src), hi = shfl_u32((uint32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}
__device__ inline uint32_t shfl_up_u32(uint32_t v, int d) { return (uint32_t)__shfl_up((int)v, d, kWave); }
__device__ inline uint64_t shfl_up_u64(uint64_t v, int d) {
    uint32_t lo = shfl_up_u32((uint32_t)v, d), hi = shfl_up_u32((uint32_t)(v >> 32), d);
    return ((uint64_t)hi << 32) | lo;
}
template <typename K> __device__ inline K shfl_xor_k(K v, int m);
template <> __device__ inline uint32_t shfl_xor_k<uint32_t>(uint32_t v, int m) { return shfl_xor_u32(v, m); }
template <> __device__ inline uint64_t shfl_xor_k<uint64_t>(uint64_t v, int m) { return shfl_xor_u64(v, m); }
template <typename K> __device__ inline K shfl_k(K v, int s);
template <> __device__ inline uint32_t shfl_k<uint32_t>(uint32_t v, int s) { return shfl_u32(v, s); }
template <> __device__ inline uint64_t shfl_k<uint64_t>(uint64_t v, int s) { return shfl_u64(v, s); }
template <typename K> __device__ inline K shfl_up_k(K v, int d);
template <> __device__ inline uint32_t shfl_up_k<uint32_t>(uint32_t v, int d) { return shfl_up_u32(v, d); }
template <> __device__ inline uint64_t shfl_up_k<uint64_t>(uint64_t v, int d) { return shfl_up_u64(v, d); }

__device__ inline uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v += shfl_xor_u64(v, m);
    return v;
}
__device__ inline uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v += shfl_xor_u32(v, m);
    return v;
}
__device__ inline uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) { uint32_t o = shfl_xor_u32(v, m); v = o < v ? o : v; }
    return v;
}
__device__ inline uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) { uint32_t o = shfl_xor_u32(v, m); v = o > v ? o : v; }
    return v;
}
// inclusive prefix sum across the wave
// Wave-wide inclusive scans by DPP moves: row shifts by 1, 2, 4 and 8 inside
// rows of 16 lanes, then lane 15 of a row into the next row (rows 1 and 3)
// and lane 31 into rows 2 and 3 -- register-to-register moves with no LDS
// round trip.  `op` combines a lane's value with one from a lower lane; a
// lane with no source combines with 0 (the identity of add, or and unsigned
// max).  Every lane of the wave must be active.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kCtrl, kRowMask, 0xF, false);
}
template <int kCtrl, int kRowMask>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v) {
    return ((uint64_t)dpp_u32<kCtrl, kRowMask>((uint32_t)(v >> 32)) << 32) |
           dpp_u32<kCtrl, kRowMask>((uint32_t)v);
}
template <int kCtrl, int kRowMask, typename T>
__device__ __forceinline__ T dpp_t(T v) {
    if constexpr (sizeof(T) == 8) return dpp_u64<kCtrl, kRowMask>(v);
    else return dpp_u32<kCtrl, kRowMask>(v);
}
template <typename T, typename Op>
__device__ __forceinline__ T wave_incl_scan_dpp(T v, Op op) {
    v = op(v, dpp_t<0x111, 0xF>(v));   // row_shr:1
    v = op(v, dpp_t<0x112, 0xF>(v));   // row_shr:2
    v = op(v, dpp_t<0x114, 0xF>(v));   // row_shr:4
    v = op(v, dpp_t<0x118, 0xF>(v));   // row_shr:8
    v = op(v, dpp_t<0x142, 0xA>(v));   // row_bcast:15 into rows 1 and 3
    v = op(v, dpp_t<0x143, 0xC>(v));   // row_bcast:31 into rows 2 and 3
    return v;
}
// Lane l - 1's value in lane l (lane 0 reads 0): wave_shr:1.
__device__ __forceinline__ uint32_t wave_shr1_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xF, 0xF, false);
}
__device__ __forceinline__ uint64_t dpp_shr1_u64(uint64_t v) {
    return ((uint64_t)wave_shr1_u32((uint32_t)(v >> 32)) << 32) | wave_shr1_u32((uint32_t)v);
}
// A lane's value read by every lane (l uniform).
__device__ __forceinline__ uint32_t readlane_u32(uint32_t v, uint32_t l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}
__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, uint32_t l) {
    return ((uint64_t)readlane_u32((uint32_t)(v >> 32), l) << 32) | readlane_u32((uint32_t)v, l);
}

__device__ inline uint32_t wave_incl_scan_u32(uint32_t v) {
    const int l = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        uint32_t o = shfl_up_u32(v, d);
        if (l >= d) v += o;
    }
    return v;
}

__device__ inline uint64_t wave_incl_scan_u64(uint64_t v) {
    const int l = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        uint64_t o = shfl_up_u64(v, d);
        if (l >= d) v += o;
    }
    return v;
}

// ---------------------------------------------------------------------------
// Register bitonic sort of 64*K keys held K per lane, element e = i*64 + lane
// (slot i).  Exchanges at distance >= 64 stay inside a lane; shorter ones use
// cross-lane shuffles.  Ascending when base = 0; for a chunk of a larger
// bitonic sort, base is the chunk's first global index and the final
// direction follows bit 64*K of base (alternating chunks).
// ---------------------------------------------------------------------------
template <int K, typename Key>
__device__ inline void wave_sort(Key (&v)[K], int base = 0) {
    const int lane = lane_id();
#pragma unroll
    for (int k = 2; k <= kWave * K; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= kWave) {
                const int js = j / kWave;
#pragma unroll
                for (int i = 0; i < K; ++i) {
                    const int ip = i ^ js;
                    if (ip > i) {
                        const bool asc = (((base + i * kWave) & k) == 0);
                        const Key a = v[i], b = v[ip];
                        const bool sw = asc ? (b < a) : (a < b);
                        v[i] = sw ? b : a;
                        v[ip] = sw ? a : b;
                    }
                }
            } else {
#pragma unroll
                for (int i = 0; i < K; ++i) {
                    const Key p = shfl_xor_k<Key>(v[i], j);
                    const int e = base + i * kWave + lane;
                    const bool asc = ((e & k) == 0);
                    const bool lower = ((lane & j) == 0);
                    const Key mn = p < v[i] ? p : v[i];
                    const Key mx = p < v[i] ? v[i] : p;
                    v[i] = (lower == asc) ? mn : mx;
                }
            }
        }
    }
}

// Bitonic merge stages j < 64*K of a sequence whose blocks of 64*K are
// bitonic with direction given by bit `k` of the global element index
// (base + e).  Used after an LDS stage of a larger sort.
template <int K, typename Key>
__device__ inline void wave_merge_tail(Key (&v)[K], int k, int base) {
    const int lane = lane_id();
#pragma unroll
    for (int j = (kWave * K) >> 1; j > 0; j >>= 1) {
        if (j >= kWave) {
            const int js = j / kWave;
#pragma unroll
            for (int i = 0; i < K; ++i) {
                const int ip = i ^ js;
                if (ip > i) {
                    const bool asc = (((base + i * kWave) & k) == 0);
                    const Key a = v[i], b = v[ip];
                    const bool sw = asc ? (b < a) : (a < b);
                    v[i] = sw ? b : a;
                    v[ip] = sw ? a : b;
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < K; ++i) {
                const Key p = shfl_xor_k<Key>(v[i], j);
                const int e = base + i * kWave + lane;
                const bool asc = ((e & k) == 0);
                const bool lower = ((lane & j) == 0);
                const Key mn = p < v[i] ? p : v[i];
                const Key mx = p < v[i] ? v[i] : p;
                v[i] = (lower == asc) ? mn : mx;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Workgroup exclusive scan of a[0..n) in LDS, in place.  Returns the total to
// every thread.  scratch: >= 17 uint32 of LDS.  Must be called by all threads.
// ---------------------------------------------------------------------------
__device__ inline uint32_t block_excl_scan(uint32_t* a, int n, uint32_t* scratch) {
    const int tid = tid_x(), nt = blockDim.x;
    const int chunk = (n + nt - 1) / nt;
    const int b = tid * chunk;
    const int e = min(b + chunk, n);
    uint32_t s = 0;
    for (int i = b; i < e; ++i) s += a[i];
    const uint32_t incl = wave_incl_scan_u32(s);
    const int w = tid / kWave, nw = nt / kWave;
    if (lane_id() == kWave - 1) scratch[w] = incl;
    __syncthreads();
    if (tid < kWave) {
        uint32_t x = tid < nw ? scratch[tid] : 0u;
        uint32_t xi = wave_incl_scan_u32(x);
        if (tid < nw) scratch[tid] = xi - x;
        if (tid == nw - 1) scratch[16] = xi;   // total (nw <= 16)
    }
    __syncthreads();
    uint32_t run = scratch[w] + incl - s;
    for (int i = b; i < e; ++i) { uint32_t t = a[i]; a[i] = run; run += t; }
    const uint32_t total = scratch[16];
    __syncthreads();
    return total;
}

// Workgroup sum of a u64 value (all threads receive it).  scratch: >= 16 u64.
__device__ inline uint64_t block_sum_u64(uint64_t v, uint64_t* scratch) {
    const int tid = tid_x(), nw = blockDim.x / kWave;
    v = wave_sum_u64(v);
    if (lane_id() == 0) scratch[tid / kWave] = v;
    __syncthreads();
    uint64_t t = 0;
    for (int i = 0; i < nw; ++i) t += scratch[i];
    __syncthreads();
    return t;
}

// ---------------------------------------------------------------------------
// Asymptotic p-value, precise path (src/cpp/mannwhitneyu.cpp:20-96), FP64 with
// contraction off so every operation rounds as the reference's does.
// alt: 0 less, 1 greater, 2 two-sided.
// ---------------------------------------------------------------------------
__device__ inline double p_asymptotic(double U1, double n1, double n2, double tie, double cc,
                                      int alt) {
#pragma clang fp contract(off)
    const double N = n1 + n2;
    const double nm = n1 * n2;
    const double mu = 0.5 * nm;
    double tie_term = 0.0;
    const double denom = N * (N - 1.0);
    if (denom > 0.0) tie_term = tie / denom;
    const double var = nm * ((N + 1.0) - tie_term) / 12.0;
    const double sigma = (var > 0.0) ? sqrt(var) : 0.0;
    double p = 1.0;
    if (alt == 2) {
        const double U2 = nm - U1;
        const double U = (U1 > U2 ? U1 : U2);
        if (sigma > 0.0) {
            const double z = (U - mu - cc) / sigma;
            p = 2.0 * (0.5 * erfc(z * 0.70710678118654752440));
            if (p > 1.0) p = 1.0;
            if (p < 0.0) p = 0.0;
        }
        return p;
    }
    const double U = (alt == 1) ? U1 : (nm - U1);
    if (sigma > 0.0) {
        const double z = (U - mu - cc) / sigma;
        p = 0.5 * erfc(z * 0.70710678118654752440);
    }
    return (p < 0.0 ? 0.0 : (p > 1.0 ? 1.0 : p));
}

}  // namespace hpdex
