// sim_runtime.h -- host execution model for the gfx950 kernel source.
//
// One OS thread per lane; a workgroup runs at a time.  Waves are 64
// consecutive threads; cross-lane operations (shuffles, ballot, all,
// readfirstlane) meet at a per-wave barrier, so every lane of a wave must
// reach them together -- exactly the requirement on the device.  A lane that
// does not arrive (a wave op under divergent control flow, or a barrier some
// threads skip) shows up as a stall, reported by a watchdog with the last
// operation each thread reached.  LDS is one host buffer; every LDS and
// global access the kernel makes through the bounds-checked helpers below is
// checked.
#pragma once
#include <algorithm>
#include <atomic>
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <vector>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __align__(n) alignas(n)
#define __shared__ static

using hipError_t = int;
using hipStream_t = void*;
using hipEvent_t = void*;
constexpr int hipSuccess = 0;
constexpr int hipErrorInvalidValue = 1;
struct uint4 { unsigned x, y, z, w; };
struct uint2 { unsigned x, y; };
inline uint2 make_uint2(unsigned a, unsigned b) { return uint2{a, b}; }

struct SimDim { unsigned x = 0, y = 0, z = 0; };
extern thread_local SimDim threadIdx;
extern SimDim blockIdx, blockDim, gridDim;

namespace sim {
struct Wave {
    std::barrier<> bar{64};
    uint64_t slot[64];
};
extern std::vector<std::unique_ptr<Wave>> g_waves;
extern std::unique_ptr<std::barrier<>> g_block_bar;
extern unsigned char* g_lds;
extern size_t g_lds_bytes;
extern thread_local const char* t_last_op;
extern unsigned char g_static_lds[4096];
inline void* static_lds(size_t bytes) {
    if (bytes > sizeof(g_static_lds)) std::abort();
    return g_static_lds;
}
inline Wave& wave() { return *g_waves[threadIdx.x / 64]; }
inline unsigned lane() { return threadIdx.x % 64; }
template <typename T> inline T exchange(T v, unsigned src) {
    Wave& w = wave();
    uint64_t bits = 0;
    std::memcpy(&bits, &v, sizeof(T));
    w.slot[lane()] = bits;
    w.bar.arrive_and_wait();
    uint64_t r = w.slot[src & 63];
    w.bar.arrive_and_wait();
    T out;
    std::memcpy(&out, &r, sizeof(T));
    return out;
}
}  // namespace sim

inline void __syncthreads() {
    sim::t_last_op = "__syncthreads";
    std::atomic_thread_fence(std::memory_order_seq_cst);
    sim::g_block_bar->arrive_and_wait();
    std::atomic_thread_fence(std::memory_order_seq_cst);
}
inline void __threadfence_block() { std::atomic_thread_fence(std::memory_order_seq_cst); }
inline void __threadfence() { std::atomic_thread_fence(std::memory_order_seq_cst); }
inline int __lane_id() { return (int)sim::lane(); }

inline int __shfl(int v, int src, int = 64) { sim::t_last_op = "shfl"; return sim::exchange(v, (unsigned)src); }
inline double __shfl(double v, int src, int = 64) { sim::t_last_op = "shfl"; return sim::exchange(v, (unsigned)src); }
inline int __shfl_xor(int v, int m, int = 64) { sim::t_last_op = "shfl_xor"; return sim::exchange(v, sim::lane() ^ (unsigned)m); }
inline int __shfl_up(int v, int d, int = 64) {
    sim::t_last_op = "shfl_up";
    const unsigned l = sim::lane();
    return sim::exchange(v, l >= (unsigned)d ? l - d : l);
}
inline unsigned long long __ballot(int p) {
    sim::t_last_op = "ballot";
    sim::Wave& w = sim::wave();
    w.slot[sim::lane()] = p ? 1 : 0;
    w.bar.arrive_and_wait();
    unsigned long long m = 0;
    for (int i = 0; i < 64; ++i) m |= (w.slot[i] ? 1ull : 0ull) << i;
    w.bar.arrive_and_wait();
    return m;
}
inline int __all(int p) { return __ballot(p) == ~0ull; }
// v_readlane: lane l's value in every lane (l uniform)
inline int __builtin_amdgcn_readlane(int v, int l) {
    sim::t_last_op = "readlane";
    return sim::exchange(v, (unsigned)l);
}
// DPP moves (gfx9 controls): row_shr:n (0x111..0x11F) within rows of 16,
// wave_shr:1 (0x138), row_bcast:15 (0x142: lane 15 of the row before) and
// row_bcast:31 (0x143: lane 31, to rows 2 and 3).  A lane of a row outside
// row_mask keeps `old`; a lane whose source is invalid reads 0 with
// bound_ctrl, else keeps `old`.
inline int __builtin_amdgcn_update_dpp(int old, int src, int ctrl, int row_mask, int bank_mask,
                                       bool bound_ctrl) {
    sim::t_last_op = "dpp";
    const unsigned l = sim::lane(), row = l / 16u, c = l % 16u;
    bool valid;
    unsigned from;
    if (ctrl >= 0x111 && ctrl <= 0x11F) {
        const unsigned n = (unsigned)ctrl - 0x110u;
        valid = c >= n;
        from = l - n;
    } else if (ctrl == 0x138) {
        valid = l >= 1u;
        from = l - 1u;
    } else if (ctrl == 0x142) {
        valid = row >= 1u;
        from = 16u * row - 1u;
    } else if (ctrl == 0x143) {
        valid = row >= 2u;
        from = 31u;
    } else {
        std::fprintf(stderr, "sim: unsupported dpp control %#x\n", ctrl);
        std::abort();
    }
    const int r = sim::exchange(src, valid ? from : l);
    const bool on = ((row_mask >> row) & 1) && ((bank_mask >> (c / 4u)) & 1);
    if (!on) return old;
    if (!valid) return bound_ctrl ? 0 : old;
    return r;
}
inline int __any(int p) { return __ballot(p) != 0ull; }
template <typename T> inline T sim_plain_load(const T* p) { return *p; }
// Buffer loads: a resource is a base address and a size in bytes; a load
// reads base + voffset + soffset, and returns 0 without reading when that
// lies past the size (the hardware range check of a raw buffer).  The
// simulator counts such loads, and rejects a nonzero scalar offset on them
// (whether the hardware adds it to the checked offset is not relied on).
struct __amdgpu_buffer_rsrc_t {
    const unsigned char* base;
    unsigned long long bytes;
};
inline __amdgpu_buffer_rsrc_t __builtin_amdgcn_make_buffer_rsrc(void* p, short, unsigned bytes, int) {
    return {(const unsigned char*)p, bytes};
}
template <typename T> inline T sim_buffer_load(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    const unsigned long long off = (unsigned long long)(unsigned)voff + (unsigned)soff;
    if (off + sizeof(T) > r.bytes) {
        if (soff != 0) {
            std::fprintf(stderr, "out-of-range buffer load with a scalar offset (%d)\n", soff);
            std::abort();
        }
        return T(0);
    }
    T v;
    std::memcpy(&v, r.base + off, sizeof(T));
    return v;
}
inline unsigned __builtin_amdgcn_raw_buffer_load_b32(__amdgpu_buffer_rsrc_t r, int voff, int soff, int) {
    return sim_buffer_load<unsigned>(r, voff, soff);
}
inline unsigned short __builtin_amdgcn_raw_buffer_load_b16(__amdgpu_buffer_rsrc_t r, int voff, int soff, int) {
    return sim_buffer_load<unsigned short>(r, voff, soff);
}
inline unsigned long long __builtin_amdgcn_raw_buffer_load_b64(__amdgpu_buffer_rsrc_t r, int voff, int soff, int) {
    return sim_buffer_load<unsigned long long>(r, voff, soff);
}
inline void __builtin_amdgcn_sched_barrier(int) {}
inline int __builtin_amdgcn_readfirstlane(int v) {
    sim::t_last_op = "readfirstlane";
    const int first = sim::exchange(v, 0u);
    if (first != v) {
        std::fprintf(stderr, "sim: readfirstlane of a non-uniform value (thread %u)\n", threadIdx.x);
        std::abort();
    }
    return first;
}
inline int __clzll(unsigned long long x) { return x ? __builtin_clzll(x) : 64; }
inline int __clz(unsigned x) { return x ? __builtin_clz(x) : 32; }
inline unsigned __float_as_uint(float f) { unsigned u; std::memcpy(&u, &f, 4); return u; }
inline long long __double_as_longlong(double d) { long long u; std::memcpy(&u, &d, 8); return u; }
inline float __uint_as_float(unsigned u) { float f; std::memcpy(&f, &u, 4); return f; }
// low 32 bits of the product of the low 24 bits of each operand (v_mul_u32_u24)
inline unsigned __umul24(unsigned a, unsigned b) { return (a & 0xFFFFFFu) * (b & 0xFFFFFFu); }
inline double __longlong_as_double(long long u) { double d; std::memcpy(&d, &u, 8); return d; }
using std::min;
using std::max;
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline unsigned min(unsigned a, unsigned b) { return a < b ? a : b; }

// atomics (host memory shared by all simulated lanes)
template <typename T> inline T sim_atomic_add(T* p, T v) { return std::atomic_ref<T>(*p).fetch_add(v); }
inline unsigned atomicAdd(unsigned* p, unsigned v) { return sim_atomic_add(p, v); }
inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v) { return sim_atomic_add(p, v); }
// compare and swap: the old value (the swap happened when it equals cmp)
inline unsigned long long atomicCAS(unsigned long long* p, unsigned long long cmp,
                                    unsigned long long v) {
    std::atomic_ref<unsigned long long>(*p).compare_exchange_strong(cmp, v);
    return cmp;
}
inline double atomicAdd(double* p, double v) {
    std::atomic_ref<double> a(*p);
    double old = a.load();
    while (!a.compare_exchange_weak(old, old + v)) {}
    return old;
}
template <typename T> inline T sim_atomic_min(T* p, T v) {
    std::atomic_ref<T> a(*p);
    T old = a.load();
    while (v < old && !a.compare_exchange_weak(old, v)) {}
    return old;
}
template <typename T> inline T sim_atomic_max(T* p, T v) {
    std::atomic_ref<T> a(*p);
    T old = a.load();
    while (v > old && !a.compare_exchange_weak(old, v)) {}
    return old;
}
inline int atomicMin(int* p, int v) { return sim_atomic_min(p, v); }
inline int atomicMax(int* p, int v) { return sim_atomic_max(p, v); }
inline unsigned atomicMax(unsigned* p, unsigned v) { return sim_atomic_max(p, v); }
inline unsigned long long atomicMin(unsigned long long* p, unsigned long long v) { return sim_atomic_min(p, v); }
inline unsigned long long atomicMax(unsigned long long* p, unsigned long long v) { return sim_atomic_max(p, v); }
inline unsigned atomicAnd(unsigned* p, unsigned v) { return std::atomic_ref<unsigned>(*p).fetch_and(v); }
inline unsigned atomicOr(unsigned* p, unsigned v) { return std::atomic_ref<unsigned>(*p).fetch_or(v); }
inline unsigned atomicExch(unsigned* p, unsigned v) { return std::atomic_ref<unsigned>(*p).exchange(v); }
#define __HIP_MEMORY_SCOPE_AGENT 3
template <typename T> inline T __hip_atomic_load(const T* p, int, int) {
    return std::atomic_ref<T>(*const_cast<T*>(p)).load();
}
