// hpdex_kernels.hip -- gfx950 kernels of the Mann-Whitney U backend.
//
// What the reference does per gene column (src/cpp/mannwhitneyu.cpp:903-1043)
// runs here as one of two launch paths over the columns of a job:
//
//   count      integer-valued columns (UMI counts): one workgroup (or one per
//              slice of the targets) reads the column from HBM once and
//              counts every entry into a (group row, value bin) counter in
//              LDS; a scan of the reference row gives lt/eq per bin, per-
//              target sums over the bins give 2U and the tie term, then U1,
//              the tie term and the asymptotic p-value are written
//                                          (group split, ranks, U, p-value)
//
//   split      every other column:
//              count stored values per group, exclusive scan, scatter each
//              group's nonzero values as order-preserving keys (or histogram
//              bins) into a per-column key range in HBM    (group split)
//   reference  sort the reference slice once per gene (on chip by buckets,
//              or register bitonic runs + LDS merges when buckets are full;
//              in HBM by merge passes beyond the on-chip size) and build its
//              rank table -- a bucket index over the sorted keys, or the
//              histogram prefix -- in HBM                            (sort)
//   target     load the column's rank table into LDS; a wave ranks windows
//              of small targets (at most 64 keys together) and targets of up
//              to 192 keys without sorting them, tie positions from an LDS
//              hash table; larger targets (and 8-byte keys) are sorted in
//              registers by a second pass; U1, sizes and the asymptotic p
//              are written                             (ranks, U, p-value)
//
// Ranking never materialises a merge: a target value v needs only
// lt(v) = #reference values below v and eq(v) = #reference values equal to v:
//
//   2*U_target = sum_b (2*lt(b) + eq(b))            over all target values b
//   U1         = n1*n2 - U_target                    (reference-group U, :960)
//   tie        = sum over distinct values of t^3 - t, built per element as
//                3*x*(x-1) with x = eq(b) + j(b), j(b) = 1-based position of b
//                inside its run of equal target values, plus the reference's
//                own runs and the zero block
//
// since f(a+b) - f(a) = sum_{j=1..b} 3(a+j)(a+j-1) for f(t) = t^3 - t.  It is
// exact integer arithmetic (2U and the tie term in 64 bits), which makes U1 and
// the tie term bit-identical to the reference's midrank merge (:618-840).
//
// Zero handling (:935-1039): zeros are never stored.  Explicit zeros, and in
// min/max/mix also absent and non-finite entries, form one block placed below
// every value (min), above every value (max) or at 0 (mix, and none -- where
// only explicit zeros count).
//
// The reference and target kernels use 256-thread workgroups and about 48 KB
// of LDS; the split kernel 1024 threads, one workgroup per CU, with up to
// 152 KB of LDS to gather a column's keys; the count kernel 256 threads and a
// tuned LDS size (40 KB, four workgroups per CU).
#include <climits>
#include <type_traits>

#include "hpdex_device.h"
#include "hpdex_internal.h"

namespace hpdex {

// Pointers taken from a JobDesc are generic; reading them as global-memory
// pointers gives global_load/global_store instead of flat accesses, whose
// completion the LDS counter (lgkmcnt) would otherwise also wait for.
#define HPDEX_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ HPDEX_GLOBAL T* gl(T* p) { return (HPDEX_GLOBAL T*)p; }
// The same for LDS: a pointer known to be in LDS gives ds_read/ds_write
// instead of flat accesses that wait on both counters.
#define HPDEX_LDS __attribute__((address_space(3)))
template <typename T>
__device__ __forceinline__ HPDEX_LDS T* lds(T* p) { return (HPDEX_LDS T*)p; }

// Buffer loads (gfx9 family): word 3 of a raw buffer resource, and the
// cache-policy bits of the loads (default policy: the non-temporal hint made
// the count kernel's column reads slower, 0.328 -> 0.307 ms without it)
constexpr int kBufferWord3 = 0x00020000;
constexpr int kBufferNT = 0;
__device__ __forceinline__ uint32_t buffer_load_u32(__amdgpu_buffer_rsrc_t rs, uint32_t voff,
                                                    uint32_t soff) {
    return __builtin_amdgcn_raw_buffer_load_b32(rs, (int)voff, (int)soff, kBufferNT);
}
__device__ __forceinline__ uint64_t buffer_load_u64(__amdgpu_buffer_rsrc_t rs, uint32_t voff) {
    return __builtin_bit_cast(uint64_t, __builtin_amdgcn_raw_buffer_load_b64(rs, (int)voff, 0, 0));
}
template <typename V>
__device__ __forceinline__ V buffer_load_value(__amdgpu_buffer_rsrc_t rs, uint32_t voff,
                                               uint32_t soff);
template <>
__device__ __forceinline__ float buffer_load_value<float>(__amdgpu_buffer_rsrc_t rs, uint32_t voff,
                                                          uint32_t soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, (int)voff, (int)soff,
                                                                          kBufferNT));
}
template <>
__device__ __forceinline__ double buffer_load_value<double>(__amdgpu_buffer_rsrc_t rs,
                                                            uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, (int)voff, (int)soff,
                                                                           kBufferNT));
}


constexpr int kChunkK = 8;            // keys per lane in a register sort
constexpr int kChunk = 64 * kChunkK;  // longest run sorted in registers

__device__ inline uint64_t tie_f(uint64_t t) { return t * t * t - t; }

// Visit the in-group entries of CSC range [p0, p1).  Each thread keeps
// kUnroll row/value loads and kUnroll group_id gathers in flight so the
// stream is not bound by one load latency per entry.
constexpr int kUnroll = 8;
template <typename V, typename F>
__device__ __forceinline__ void for_each_entry(const int32_t* __restrict__ rows,
                                               const V* __restrict__ vals,
                                               const int32_t* __restrict__ gid, uint32_t R,
                                               uint32_t G, int64_t p0, int64_t p1, F&& f) {
    const int tid = tid_x(), nt = blockDim.x;
    for (int64_t pb = p0; pb < p1; pb += (int64_t)nt * kUnroll) {
        uint32_t r[kUnroll];
        V v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const int64_t p = pb + (int64_t)u * nt + tid;
            if (p < p1) { r[u] = (uint32_t)gl(rows)[p]; v[u] = gl(vals)[p]; }
            else { r[u] = 0xFFFFFFFFu; v[u] = V(0); }
        }
        uint32_t g[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) g[u] = r[u] < R ? (uint32_t)gl(gid)[r[u]] : 0xFFFFFFFFu;
#pragma unroll
        for (int u = 0; u < kUnroll; ++u)
            if (g[u] < G) f(g[u], v[u]);
    }
}

// Visit (group, value, entry index in the column) of every entry of the
// column [p0, p0 + n) whose row is in a group, for workgroups of NT threads
// (kBatch: f takes a whole batch, kUnroll entries per thread, entries outside
// every group with a group of G or above).  Batches of NT*kUnroll entries are
// read as buffer loads whose resources end at the batch's last entry (lanes
// past it read zeros and are masked) in a two-stage software pipeline: while
// batch b is visited, the labels of batch b + 1 are being gathered and the
// rows and values of batch b + 2 loaded (loads complete in order: waiting for
// one batch leaves the later ones outstanding).  kPacked gathers the labels
// from the bit-packed array (rows out of range read its sentinel at index R):
// a sparse column's gathers then touch as few cache lines as possible.
// Otherwise the 32-bit labels are gathered: for a column holding a large
// share of the rows they are coalesced and unpacking would only add work.
template <typename V, int NT, bool kPacked, bool kBatch = false, int U = kUnroll, typename F>
__device__ __forceinline__ void for_each_grouped_pipelined(const JobDesc& J, int64_t p0, int64_t n,
                                                           F&& f) {
    const uint32_t tid = (uint32_t)tid_x();
    const uint32_t R = (uint32_t)J.R, G = (uint32_t)J.G;
    const uint32_t lbits = J.gid_bits, lmask = (1u << lbits) - 1u;
    const __amdgpu_buffer_rsrc_t gid_rsrc = __builtin_amdgcn_make_buffer_rsrc(
        (void*)J.gidp, (short)0, kPacked ? J.gid_words * 4u : 0u, kBufferWord3);
    const int32_t* col_rows = J.rows + p0;
    const V* col_vals = (const V*)J.vals + p0;
    constexpr uint32_t step = (uint32_t)NT * U;
    // rows and values of the batch at ib
    auto load = [&](int64_t ib, uint32_t (&r)[U], V (&v)[U]) {
        const int64_t left = n - ib;
        const uint32_t m = left < (int64_t)step ? (uint32_t)left : step;
        const __amdgpu_buffer_rsrc_t rows_rsrc = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(col_rows + ib), (short)0, m * 4u, kBufferWord3);
        const __amdgpu_buffer_rsrc_t vals_rsrc = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(col_vals + ib), (short)0, m * (uint32_t)sizeof(V), kBufferWord3);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t e = tid + (uint32_t)(u * NT);
            r[u] = buffer_load_u32(rows_rsrc, e * 4u, 0u);
            v[u] = buffer_load_value<V>(vals_rsrc, e * (uint32_t)sizeof(V), 0u);
        }
    };
    // labels of the batch at ib; lanes past the column's end take none
    auto gather = [&](int64_t ib, const uint32_t (&r)[U], uint32_t (&g)[U]) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool in = ib + (int64_t)(tid + (uint32_t)(u * NT)) < n;
            const uint32_t ru = in ? r[u] : 0xFFFFFFFFu;
            if (kPacked) {
                const uint32_t pos = min(ru, R) * lbits;
                const uint64_t w2 = buffer_load_u64(gid_rsrc, (pos >> 5) * 4u);
                g[u] = (uint32_t)(w2 >> (pos & 31u)) & lmask;
            } else {
                g[u] = ru < R ? (uint32_t)gl(J.gid)[ru] : 0xFFFFFFFFu;
            }
        }
    };
    auto visit = [&](int64_t ib, const uint32_t (&g)[U], const V (&v)[U]) {
        if constexpr (kBatch) {
            f(g, v, (uint32_t)ib + tid);
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (g[u] < G) f(g[u], v[u], (uint32_t)ib + tid + (uint32_t)(u * NT));
        }
    };
    if (n <= 0) return;
    uint32_t r[U], g[U];
    V v[U];
    load(0, r, v);
    gather(0, r, g);
    int64_t ib = 0;
    if ((int64_t)step < n) {
        uint32_t r2[U];
        V v2[U];
        load(step, r2, v2);
        for (;;) {   // batch ib with the batch at ib + step after it
            uint32_t g2[U];
            gather(ib + step, r2, g2);
            const bool more = ib + 2 * (int64_t)step < n;
            uint32_t r3[U];
            V v3[U];
            if (more) load(ib + 2 * step, r3, v3);
            __builtin_amdgcn_sched_barrier(0);
            visit(ib, g, v);
            ib += step;
#pragma unroll
            for (int u = 0; u < U; ++u) { g[u] = g2[u]; v[u] = v2[u]; }
            if (!more) break;
#pragma unroll
            for (int u = 0; u < U; ++u) { r2[u] = r3[u]; v2[u] = v3[u]; }
        }
    }
    visit(ib, g, v);   // the column's last batch
}

// The one-stage reader: the next batch's rows and values are issued after
// this batch's label gathers and before its visits.  kInOrder (the column
// stores rows 0 .. R-1 in order): an entry's row is its position, only its
// value is read; its label gathers are coalesced (for such a column, labels
// loaded with the values one stage earlier, in smaller batches, were slower:
// 1.64 -> 1.89 ms on a dense 150,000 x 6,000 split).
template <typename V, int NT, bool kPacked, bool kInOrder = false, bool kBatch = false,
          int U = kUnroll, typename F>
__device__ __forceinline__ void for_each_grouped(const JobDesc& J, int64_t p0, int64_t n, F&& f) {
    const uint32_t tid = (uint32_t)tid_x();
    const uint32_t R = (uint32_t)J.R, G = (uint32_t)J.G;
    const uint32_t lbits = J.gid_bits, lmask = (1u << lbits) - 1u;
    const __amdgpu_buffer_rsrc_t gid_rsrc = __builtin_amdgcn_make_buffer_rsrc(
        (void*)J.gidp, (short)0, kPacked ? J.gid_words * 4u : 0u, kBufferWord3);
    const int32_t* col_rows = J.rows + p0;
    const V* col_vals = (const V*)J.vals + p0;
    constexpr uint32_t step = (uint32_t)NT * U;
    auto load = [&](int64_t ib, uint32_t (&r)[U], V (&v)[U]) {
        const int64_t left = n - ib;
        const uint32_t m = left < (int64_t)step ? (uint32_t)left : step;
        const __amdgpu_buffer_rsrc_t rows_rsrc = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(col_rows + ib), (short)0, m * 4u, kBufferWord3);
        const __amdgpu_buffer_rsrc_t vals_rsrc = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(col_vals + ib), (short)0, m * (uint32_t)sizeof(V), kBufferWord3);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t e = tid + (uint32_t)(u * NT);
            // kInOrder: the column stores rows 0 .. R-1 in order, so an
            // entry's row is its position (R < 2^31 keeps it in 32 bits)
            r[u] = kInOrder ? (uint32_t)ib + e : buffer_load_u32(rows_rsrc, e * 4u, 0u);
            v[u] = buffer_load_value<V>(vals_rsrc, e * (uint32_t)sizeof(V), 0u);
        }
    };
    auto gather = [&](const uint32_t (&r)[U], uint32_t (&g)[U]) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (kPacked) {
                const uint32_t pos = min(r[u], R) * lbits;
                const uint64_t w2 = buffer_load_u64(gid_rsrc, (pos >> 5) * 4u);
                g[u] = (uint32_t)(w2 >> (pos & 31u)) & lmask;
            } else {
                g[u] = r[u] < R ? (uint32_t)gl(J.gid)[r[u]] : 0xFFFFFFFFu;
            }
        }
    };
    if (n <= 0) return;
    uint32_t r[U];
    V v[U];
    load(0, r, v);
    int64_t ib = 0;
    for (; ib + (int64_t)step < n; ib += step) {   // a whole batch with another after it
        uint32_t g[U];
        gather(r, g);
        __builtin_amdgcn_sched_barrier(0);
        uint32_t r2[U];
        V v2[U];
        load(ib + step, r2, v2);
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (kBatch) {
            f(g, v, (uint32_t)ib + tid);
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (g[u] < G) f(g[u], v[u], (uint32_t)ib + tid + (uint32_t)(u * NT));
        }
#pragma unroll
        for (int u = 0; u < U; ++u) { r[u] = r2[u]; v[u] = v2[u]; }
    }
    // the last batch: lanes past the column's end masked
    const uint32_t m = (uint32_t)(n - ib);
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (tid + (uint32_t)(u * NT) >= m) r[u] = 0xFFFFFFFFu;
    uint32_t g[U];
    gather(r, g);
    if constexpr (kBatch) {
        f(g, v, (uint32_t)ib + tid);
    } else {
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (g[u] < G) f(g[u], v[u], (uint32_t)ib + tid + (uint32_t)(u * NT));
    }
}


__device__ inline void raise_status(const JobDesc& J, uint32_t bit, int64_t c) {
    atomicOr(J.status, bit);
    atomicMin(J.status_col, (unsigned long long)c);
}

template <typename K>
__device__ inline K wave_min_k(K v) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) { K o = shfl_xor_k<K>(v, m); v = o < v ? o : v; }
    return v;
}
template <typename K>
__device__ inline K wave_max_k(K v) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) { K o = shfl_xor_k<K>(v, m); v = o > v ? o : v; }
    return v;
}
__device__ inline int wave_min_i(int v) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v = min(v, __shfl_xor(v, m, kWave));
    return v;
}
__device__ inline int wave_max_i(int v) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v = max(v, __shfl_xor(v, m, kWave));
    return v;
}

// first index in a[lo, hi) with a[i] >= v
template <typename K, typename P>
__device__ inline uint32_t lower_bound_k(P a, uint32_t lo, uint32_t hi, K v) {
    uint32_t n = hi - lo;
    while (n > 0) {
        const uint32_t h = n >> 1;
        if (a[lo + h] < v) { lo += h + 1; n -= h + 1; } else { n = h; }
    }
    return lo;
}
// first index in a[lo, hi) with a[i] > v
template <typename K, typename P>
__device__ inline uint32_t upper_bound_k(P a, uint32_t lo, uint32_t hi, K v) {
    uint32_t n = hi - lo;
    while (n > 0) {
        const uint32_t h = n >> 1;
        if (!(v < a[lo + h])) { lo += h + 1; n -= h + 1; } else { n = h; }
    }
    return lo;
}

// The work item of this workgroup, uniform across it (kept in SGPRs).
__device__ inline void work_item(const WorkItem* items, uint32_t k, int& job, int64_t& col) {
    job = __builtin_amdgcn_readfirstlane(items[k].job);
    col = (int64_t)__builtin_amdgcn_readfirstlane(items[k].col);
}

template <typename T> struct TypeTag { using type = T; };

// A workgroup-uniform value held in an SGPR.
__device__ __forceinline__ uint32_t uniform_u32(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)x);
}
__device__ __forceinline__ int64_t uniform_i64(int64_t x) {
    const uint64_t u = (uint64_t)x;
    const uint32_t lo = uniform_u32((uint32_t)u), hi = uniform_u32((uint32_t)(u >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Zero-block sizes and sample sizes of pair (reference, g) in column c
// (mannwhitneyu.cpp:936-1037): none counts stored finite values only; the
// other modes count every group row, absent and non-finite entries as zeros.
__device__ inline void pair_sizes(const JobDesc& J, int64_t c, uint32_t g, uint32_t& za,
                                  uint32_t& zb, uint32_t& n1, uint32_t& n2) {
    const int64_t b = c * J.G;
    const uint32_t nza = gl(J.cnt)[b], nzb = gl(J.cnt)[b + g];
    if (J.mode == kZeroNone) {
        za = gl(J.zcnt)[b]; zb = gl(J.zcnt)[b + g];
        n1 = nza + za; n2 = nzb + zb;
    } else {
        const uint32_t gr0 = gl(J.group_rows)[0], grg = gl(J.group_rows)[g];
        za = gr0 > nza ? gr0 - nza : 0u;
        zb = grg > nzb ? grg - nzb : 0u;
        n1 = gr0; n2 = grg;
    }
}

// ---------------------------------------------------------------------------
// split: per-group counts, value statistics, errors, offsets and the scatter
// of nonzero values into the column's key range (mannwhitneyu.cpp:913-926).
// ---------------------------------------------------------------------------
struct SplitState {
    uint32_t all_int;
    int32_t vmin, vmax;
    uint32_t hist;
    uint32_t scan[20];
};

// Buckets of a sort-path column's reference index: the power of two of at
// least na, 64 to kMaxBuckets.
__device__ inline uint32_t ref_buckets(uint32_t na) {
    uint32_t nbk = 64;
    while (nbk < na && nbk < (uint32_t)kMaxBuckets) nbk <<= 1;
    return nbk;
}

// bucket = (key - kmin) >> shift: the shift that maps the key range onto nb
// buckets
__device__ inline uint32_t ref_shift(uint64_t kmin, uint64_t kmax, uint32_t na, uint32_t nb) {
    const uint64_t rng = na > 0 ? kmax - kmin : 0ull;
    const uint32_t bits = rng ? 64u - (uint32_t)__clzll(rng) : 0u;
    const uint32_t lg = 31u - (uint32_t)__clz(nb);
    return bits > lg ? bits - lg : 0u;
}

// The column's plan from its split statistics (na: its reference's stored
// nonzero finite values): the histogram branch when every value is an integer
// of a narrow enough range, else the sort path with a bucket index of nb
// buckets.  The reference kernel adds the reference's key range, bucket shift
// and values below zero (it reads every reference value anyway, sorted).
// Leaves the plan's hist and vmin in st.
__device__ inline void write_column_plan(const JobDesc& J, int64_t c, uint32_t na, SplitState& st) {
    ColPlan pl;
    pl.na = na;
    pl.refneg = 0;
    pl.kmin = ~0ull;
    pl.kmax = 0;
    pl.shift = 0;
    pl.TA = 0;
    pl.maxb = 0;
    pl.bigt = 0;
    const int64_t range = (st.vmax >= st.vmin) ? (int64_t)st.vmax - st.vmin + 1 : 1;
    if (J.hist_max > 0 && st.all_int && range <= J.hist_max) {
        pl.hist = 1;
        pl.nb = (uint32_t)range;
        pl.vmin = (st.vmax >= st.vmin) ? st.vmin : 0;
    } else {
        pl.hist = 0;
        pl.nb = ref_buckets(na);
        pl.vmin = 0;
    }
    J.plan[c] = pl;
    st.hist = pl.hist;
    st.vmin = pl.vmin;
}

template <typename V, int NT>
__device__ __forceinline__ void split_item(const JobDesc* jobs, const WorkItem* items, uint32_t k,
                                           uint32_t lds_bytes) {
    using KT = KeyOf<V>;
    using K = typename KT::K;
    // entries per thread per batch of the two-stage column reader
    constexpr int kSplitUnroll = sizeof(V) == 4 ? 4 : 2;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ SplitState st;
    int job;
    int64_t c;
    work_item(items, k, job, c);
    const JobDesc J = jobs[job];
    const int tid = tid_x(), nt = blockDim.x, lane = lane_id();
    const int G = J.G;
    const uint32_t R = (uint32_t)J.R;
    const V* vals = (const V*)J.vals;
    const int64_t p0 = J.indptr[c], p1 = J.indptr[c + 1];
    if (split_counter_bytes((uint64_t)G) > (uint64_t)lds_bytes) {   // the launch takes
        if (tid == 0) raise_status(J, kStatusGroups, c);               // split_global_kernel
        return;
    }
    // per group: explicit zeros (high half) and nonzero values (low half),
    // then the offsets; kSplitDummy extra words of each take the no-op
    // counts of entries outside every group
    unsigned long long* cnt = (unsigned long long*)smem;
    uint32_t* cur = (uint32_t*)(cnt + G + kSplitDummy);
    const uint32_t dummy = (uint32_t)G + ((uint32_t)lane & (kSplitDummy - 1u));
    // the column's entries in a group: bit-packed labels for a sparse column
    // (its rows are far apart, so 32-bit label gathers would touch a cache
    // line per entry), 32-bit labels for a column holding a quarter of the
    // rows or more (consecutive rows: the gathers are coalesced)
    const bool packed = __builtin_amdgcn_readfirstlane(
        (J.gid_bits > 0u && 4 * (p1 - p0) < (int64_t)R) ? 1 : 0) != 0;
    // a column storing rows 0 .. R-1 in order (a dense matrix held as CSC):
    // rows from positions, so only its values are read
    const bool in_order = __builtin_amdgcn_readfirstlane(J.colident && J.colident[c] ? 1 : 0) != 0;
    // a column whose values are not all small integers (the matrix's metadata)
    // cannot take the histogram branch: its values' statistics are not taken
    const bool int_col = __builtin_amdgcn_readfirstlane(!J.colint || J.colint[c] ? 1 : 0) != 0;
    // sparse columns through the two-stage reader (the label gathers of one
    // batch overlap the visits of the one before: 2.06 -> 1.98 ms on a
    // 400,000 x 18,000 split at 5 %), columns in row order through the
    // one-stage reader with rows from positions
    auto grouped = [&](auto&& f) {
        if (packed) for_each_grouped_pipelined<V, NT, true, false, kSplitUnroll>(J, p0, p1 - p0, f);
        else if (in_order) for_each_grouped<V, NT, false, true, false, kUnroll>(J, p0, p1 - p0, f);
        else for_each_grouped_pipelined<V, NT, false, false, kSplitUnroll>(J, p0, p1 - p0, f);
    };
    // every entry of a batch (group G or above: none)
    auto grouped_batch = [&](auto&& f) {
        if (packed) for_each_grouped_pipelined<V, NT, true, true, kSplitUnroll>(J, p0, p1 - p0, f);
        else if (in_order) for_each_grouped<V, NT, false, true, true, kUnroll>(J, p0, p1 - p0, f);
        else for_each_grouped_pipelined<V, NT, false, true, kSplitUnroll>(J, p0, p1 - p0, f);
    };

    // Slots: group g's keys at the exclusive prefix of the group sizes.  When
    // the column's key range holds them all (a column storing at least as many
    // entries as there are grouped rows, e.g. a dense one), every stored value
    // goes to its slot in the counting pass itself and the column is read once;
    // otherwise the counts are scanned into offsets and a second pass scatters.
    for (int g = tid; g < G + (int)kSplitDummy; g += nt) {
        cnt[g] = 0ull;
        cur[g] = g < G ? J.group_rows[g] : 0u;
    }
    if (tid == 0) {
        st.all_int = 1u; st.vmin = INT_MAX; st.vmax = INT_MIN;
    }
    __syncthreads();
    const uint32_t slots = block_excl_scan(cur, G, st.scan);
    const bool fixed = __builtin_amdgcn_readfirstlane((int64_t)slots <= p1 - p0 ? 1 : 0) != 0;
    K* keys = (K*)J.keys + p0;
    // the keys gather in LDS when the workgroup has room for the column's key
    // range (its slots, or its entries when it is scattered by offsets) and
    // go to HBM afterwards in one coalesced copy: scattered 4-byte stores to
    // HBM write whole partial lines
    const uint32_t cnt_bytes = (uint32_t)split_counter_bytes((uint64_t)G);
    const uint64_t span = fixed ? (uint64_t)slots : (uint64_t)(p1 - p0);
    const bool staged = __builtin_amdgcn_readfirstlane(
        (uint64_t)cnt_bytes + (uint64_t)sizeof(K) * span <= lds_bytes ? 1 : 0) != 0;
    K* stage = (K*)(smem + cnt_bytes);
    // A column scattered by offsets whose entries fit LDS keeps each entry's
    // group (16 bits, after the key space) and key (in the key space, in
    // arrival order) from its counting pass, and the scatter reads them
    // there: no second pass over the rows, values or labels.  Entries that
    // give no key keep group 0xFFFF.
    const uint32_t n_col = (uint32_t)(p1 - p0);
    const uint32_t stage_bytes = ((uint32_t)sizeof(K) * n_col + 15u) & ~15u;
    const bool arrival = __builtin_amdgcn_readfirstlane(
        (!fixed && J.split_arrival &&
         (uint64_t)cnt_bytes + stage_bytes + 2ull * n_col <= (uint64_t)lds_bytes) ? 1 : 0) != 0;
    uint16_t* agrp = (uint16_t*)(smem + cnt_bytes + stage_bytes);
    if (arrival)
        for (uint32_t e = (uint32_t)tid; e < n_col; e += (uint32_t)nt) agrp[e] = 0xFFFFu;

    // counts and value statistics (and, in slots, the keys), a batch of
    // kUnroll entries per thread at a time with no branch around the
    // counters: every entry adds to its group's word or to a dummy word, so
    // the batch's LDS atomics are issued back to back and waited for once
    // fixed_tag: 0 scatter later, 1 keys to their slots in HBM, 2 in LDS
    auto count_pass = [&](auto fixed_tag) {
        constexpr int kFixed = decltype(fixed_tag)::value;
        bool li = int_col;
        int lmin = INT_MAX, lmax = INT_MIN;
        grouped_batch([&](const auto& g, const auto& v, uint32_t e0) {
            constexpr int U = (int)(sizeof(g) / sizeof(g[0]));
            bool nz[U];
            uint32_t gi[U], pos[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const bool in = g[u] < (uint32_t)G && KT::finite(v[u]);
                nz[u] = in && !(v[u] == V(0));
                gi[u] = in ? g[u] : dummy;
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                pos[u] = (uint32_t)atomicAdd(&cnt[gi[u]], nz[u] ? 1ull : (1ull << 32));
            if constexpr (kFixed != 0) {
#pragma unroll
                for (int u = 0; u < U; ++u) pos[u] += cur[nz[u] ? gi[u] : dummy];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (!nz[u]) continue;
                const K k = KT::key(v[u]);
                if (kFixed == 0) {
                    if (arrival) {
                        agrp[e0 + (uint32_t)(u * NT)] = (uint16_t)g[u];
                        stage[e0 + (uint32_t)(u * NT)] = k;
                    }
                } else if (kFixed == 1) {
                    // a group holding more entries than rows (repeated row
                    // indices) is an error of the column; its keys stay
                    // inside the column's range
                    if ((int64_t)pos[u] < p1 - p0) keys[pos[u]] = k;
                } else {
                    if (pos[u] < slots) stage[pos[u]] = k;
                }
                if (int_col) {
                    int iv;
                    if (KT::small_int(v[u], iv)) { lmin = min(lmin, iv); lmax = max(lmax, iv); }
                    else li = false;
                }
            }
        });
        const bool wall = __all(li ? 1 : 0);
        lmin = wave_min_i(lmin);
        lmax = wave_max_i(lmax);
        if (lane == 0) {
            if (!wall) atomicAnd(&st.all_int, 0u);
            atomicMin(&st.vmin, lmin);
            atomicMax(&st.vmax, lmax);
        }
    };
    // A column storing rows 0 .. R-1 in order (a dense matrix held as CSC):
    // the value of row r is entry r, so only the group members' values are
    // read, gathered through the job's member list (rows ascending: a wave's
    // 64 members span a few hundred rows, a few cache lines); the rows in no
    // group -- most of a dense column when the groups are small -- take no
    // label gather, counter or key.  Counts, slot keys and statistics as the
    // count pass gives them.
    auto member_pass = [&](auto fixed_tag) {
        constexpr int kFixed = decltype(fixed_tag)::value;   // 1 slots in HBM, 2 in LDS
        constexpr int U = 4;
        const uint32_t M = J.n_members;
        const V* col = vals + p0;
        bool li = int_col;
        int lmin = INT_MAX, lmax = INT_MIN;
        // the member list of the next batch loads while this batch's values
        // are gathered
        uint32_t rn[U], gn[U];
        auto members_of = [&](uint32_t s0, uint32_t (&r)[U], uint32_t (&g)[U]) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t sl = s0 + (uint32_t)(u * nt);
                const bool ok = sl < M;
                r[u] = ok ? gl(J.members)[sl] : 0u;
                g[u] = ok ? gl(J.member_group)[sl] : dummy;
            }
        };
        members_of((uint32_t)tid, rn, gn);
        for (uint32_t s0 = (uint32_t)tid; s0 < M; s0 += (uint32_t)nt * U) {
            uint32_t gi[U], pos[U];
            V v[U];
            bool nz[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                gi[u] = gn[u];
                v[u] = s0 + (uint32_t)(u * nt) < M ? gl(col)[rn[u]] : V(0);
            }
            members_of(s0 + (uint32_t)nt * U, rn, gn);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const bool in = gi[u] < (uint32_t)G && KT::finite(v[u]);
                nz[u] = in && !(v[u] == V(0));
                if (!in) gi[u] = dummy;
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                pos[u] = (uint32_t)atomicAdd(&cnt[gi[u]], nz[u] ? 1ull : (1ull << 32));
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (!nz[u]) continue;
                const uint32_t at = pos[u] + cur[gi[u]];
                const K k = KT::key(v[u]);
                if (kFixed == 1) {
                    if ((int64_t)at < p1 - p0) keys[at] = k;
                } else {
                    if (at < slots) stage[at] = k;
                }
                if (int_col) {
                    int iv;
                    if (KT::small_int(v[u], iv)) { lmin = min(lmin, iv); lmax = max(lmax, iv); }
                    else li = false;
                }
            }
        }
        const bool wall = __all(li ? 1 : 0);
        lmin = wave_min_i(lmin);
        lmax = wave_max_i(lmax);
        if (lane == 0) {
            if (!wall) atomicAnd(&st.all_int, 0u);
            atomicMin(&st.vmin, lmin);
            atomicMax(&st.vmax, lmax);
        }
    };
    const bool by_members =
        __builtin_amdgcn_readfirstlane(in_order && fixed && J.members ? 1 : 0) != 0;
    if (arrival) __syncthreads();   // the cleared groups before the pass writes them
    if (by_members && staged) member_pass(std::integral_constant<int, 2>{});
    else if (by_members) member_pass(std::integral_constant<int, 1>{});
    else if (fixed && staged) count_pass(std::integral_constant<int, 2>{});
    else if (fixed) count_pass(std::integral_constant<int, 1>{});
    else count_pass(std::integral_constant<int, 0>{});
    __syncthreads();

    // errors of this column (mannwhitneyu.cpp:936-1038), counts out
    {
        const uint64_t ebase = (uint64_t)c * (uint64_t)(2 * G + 1);
        const int64_t b = c * G;
        for (int g = tid; g < G; g += nt) {
            const uint32_t nz = (uint32_t)cnt[g], ez = (uint32_t)(cnt[g] >> 32),
                           gr = J.group_rows[g];
            if ((uint64_t)nz + ez > gr) atomicMin(J.err_key, (unsigned long long)(ebase + g));
            const uint32_t n = (J.mode == kZeroNone) ? nz + ez : gr;
            if (n < J.min_n) atomicMin(J.err_key, (unsigned long long)(ebase + (g == 0 ? G : G + g)));
            J.cnt[b + g] = nz;
            J.zcnt[b + g] = ez;
            if (!fixed) cur[g] = nz;
        }
    }
    __syncthreads();
    if (!fixed) block_excl_scan(cur, G, st.scan);
    for (int g = tid; g < G; g += nt) J.goff[c * G + g] = cur[g];

    // plan of the column
    if (tid == 0) write_column_plan(J, c, (uint32_t)cnt[0], st);
    __syncthreads();

    const uint32_t hist = st.hist;
    const int32_t vmin = st.vmin;
    if (arrival) {
        // scatter from the recorded groups and keys (LDS, in arrival order)
        // straight to the keys' places in HBM: one workgroup per CU keeps
        // few columns' key ranges in flight, so their lines fill in L2
        for (uint32_t e0 = (uint32_t)tid; e0 < n_col; e0 += (uint32_t)nt * 4u) {
            uint32_t g[4], pos[4];
            K kk[4];
#pragma unroll
            for (uint32_t u = 0; u < 4u; ++u) {
                const uint32_t e = e0 + u * (uint32_t)nt;
                g[u] = e < n_col ? (uint32_t)agrp[e] : 0xFFFFu;
                kk[u] = g[u] != 0xFFFFu ? stage[e] : K(0);
            }
#pragma unroll
            for (uint32_t u = 0; u < 4u; ++u)
                pos[u] = atomicAdd(&cur[g[u] != 0xFFFFu ? g[u] : dummy], 1u);
#pragma unroll
            for (uint32_t u = 0; u < 4u; ++u) {
                if (g[u] == 0xFFFFu) continue;
                keys[pos[u]] = hist ? (K)(uint32_t)((int)KT::value(kk[u]) - vmin) : kk[u];
            }
        }
    } else if (!fixed) {
        // scatter nonzero values into the groups' key ranges
        K* dst = staged ? stage : keys;
        grouped([&](uint32_t g, V v, uint32_t) {
            if (!KT::finite(v) || v == V(0)) return;
            const uint32_t pos = atomicAdd(&cur[g], 1u);
            dst[pos] = hist ? (K)(uint32_t)((int)v - vmin) : KT::key(v);
        });
        if (staged) {   // the column's key range out, coalesced
            __syncthreads();
            const uint32_t n = (uint32_t)(p1 - p0);
            for (uint32_t i = (uint32_t)tid; i < n; i += (uint32_t)nt) keys[i] = stage[i];
        }
    } else {
        K* slot_keys = staged ? stage : keys;
        if (hist) {
            // the slots hold keys; a histogram plan reads bins: each group's
            // keys (integers, as the statistics found them) become v - vmin
            const int w = tid / kWave, nw = nt / kWave;
            for (int g = w; g < G; g += nw) {
                const uint32_t o = cur[g], m = min((uint32_t)cnt[g], J.group_rows[g]);
                for (uint32_t i = (uint32_t)lane; i < m; i += kWave)
                    slot_keys[o + i] = (K)(uint32_t)((int)KT::value(slot_keys[o + i]) - vmin);
            }
        }
        if (staged) {   // the whole slot range out, coalesced
            __syncthreads();
            for (uint32_t i = (uint32_t)tid; i < slots; i += (uint32_t)nt) keys[i] = stage[i];
        }
    }
}

// workgroup b takes queue item b (the grid is sized for the queue's capacity).
// kSplitBlock threads, one workgroup per CU: a column's keys are scattered
// over its key range group by group, and with few columns in flight those
// ranges stay in L2 while their partial lines fill (256-thread workgroups, five
// per CU, wrote 8x the key bytes to HBM: 4.0 -> 3.6 ms on a dense 150,000 x
// 6,000 split, 3.2 -> 2.3 ms on a sparse 400,000 x 18,000 one).
template <typename V>
__global__ __launch_bounds__(kSplitBlock) void split_kernel(const JobDesc* jobs,
                                                            const WorkItem* items,
                                                            uint32_t n_items,
                                                            uint32_t lds_bytes) {
    if (blockIdx.x < n_items) split_item<V, kSplitBlock>(jobs, items, blockIdx.x, lds_bytes);
}

// A device-scope load of a counter that device-scope atomics update (it
// bypasses the CU's L1, which may hold an older copy of the line).
__device__ __forceinline__ uint32_t load_counter(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The split of a job with more groups than the workgroup's LDS holds
// counters for: the outputs of split_item, with the per-group counts kept in
// the job's cnt / zcnt arrays in HBM (device-scope atomics) and the slot
// starts or offsets scanned in its goff array; keys go straight to their
// slots or offsets in HBM.  A scatter counts cnt up from zero again as each
// group's cursor, which leaves the same counts behind.
template <typename V, int NT>
__device__ __forceinline__ void split_global_item(const JobDesc* jobs, const WorkItem* items,
                                                  uint32_t k) {
    using KT = KeyOf<V>;
    using K = typename KT::K;
    __shared__ SplitState st;
    int job;
    int64_t c;
    work_item(items, k, job, c);
    const JobDesc J = jobs[job];
    const int tid = tid_x(), nt = blockDim.x, lane = lane_id();
    const int G = J.G;
    const uint32_t R = (uint32_t)J.R;
    const int64_t p0 = J.indptr[c], p1 = J.indptr[c + 1];
    const uint32_t n_col = (uint32_t)(p1 - p0);
    uint32_t* cnt = J.cnt + c * G;
    uint32_t* zcnt = J.zcnt + c * G;
    uint32_t* off = J.goff + c * G;
    const bool packed = __builtin_amdgcn_readfirstlane(
        (J.gid_bits > 0u && 4 * (p1 - p0) < (int64_t)R) ? 1 : 0) != 0;
    const bool in_order = __builtin_amdgcn_readfirstlane(J.colident && J.colident[c] ? 1 : 0) != 0;
    const bool int_col = __builtin_amdgcn_readfirstlane(!J.colint || J.colint[c] ? 1 : 0) != 0;
    auto grouped = [&](auto&& f) {
        if (packed) for_each_grouped_pipelined<V, NT, true, false, 4>(J, p0, p1 - p0, f);
        else if (in_order) for_each_grouped<V, NT, false, true, false, kUnroll>(J, p0, p1 - p0, f);
        else for_each_grouped_pipelined<V, NT, false, false, 4>(J, p0, p1 - p0, f);
    };
    for (int g = tid; g < G; g += nt) {
        atomicExch(&cnt[g], 0u);
        atomicExch(&zcnt[g], 0u);
        off[g] = J.group_rows[g];
    }
    if (tid == 0) {
        st.all_int = 1u; st.vmin = INT_MAX; st.vmax = INT_MIN;
    }
    __syncthreads();
    // slots: group g's keys at the prefix of the group sizes when the column's
    // key range holds every group's rows
    const uint32_t slots = block_excl_scan(off, G, st.scan);
    const bool fixed = __builtin_amdgcn_readfirstlane((int64_t)slots <= p1 - p0 ? 1 : 0) != 0;
    K* keys = (K*)J.keys + p0;
    {
        bool li = int_col;
        int lmin = INT_MAX, lmax = INT_MIN;
        grouped([&](uint32_t g, V v, uint32_t) {
            if (!KT::finite(v)) return;
            if (v == V(0)) {
                atomicAdd(&zcnt[g], 1u);
                return;
            }
            const uint32_t pos = atomicAdd(&cnt[g], 1u);
            const K key = KT::key(v);
            // a group holding more entries than rows (repeated row indices)
            // is an error of the column; its keys stay inside the column's range
            if (fixed && off[g] + pos < n_col) keys[off[g] + pos] = key;
            if (int_col) {
                int iv;
                if (KT::small_int(v, iv)) { lmin = min(lmin, iv); lmax = max(lmax, iv); }
                else li = false;
            }
        });
        const bool wall = __all(li ? 1 : 0);
        lmin = wave_min_i(lmin);
        lmax = wave_max_i(lmax);
        if (lane == 0) {
            if (!wall) atomicAnd(&st.all_int, 0u);
            atomicMin(&st.vmin, lmin);
            atomicMax(&st.vmax, lmax);
        }
    }
    __syncthreads();
    // errors of this column (mannwhitneyu.cpp:936-1038); offsets from the counts
    const uint64_t ebase = (uint64_t)c * (uint64_t)(2 * G + 1);
    for (int g = tid; g < G; g += nt) {
        const uint32_t nz = load_counter(&cnt[g]), ez = load_counter(&zcnt[g]),
                       gr = J.group_rows[g];
        if ((uint64_t)nz + ez > gr) atomicMin(J.err_key, (unsigned long long)(ebase + g));
        const uint32_t n = (J.mode == kZeroNone) ? nz + ez : gr;
        if (n < J.min_n) atomicMin(J.err_key, (unsigned long long)(ebase + (g == 0 ? G : G + g)));
        if (!fixed) off[g] = nz;
    }
    __syncthreads();
    if (!fixed) block_excl_scan(off, G, st.scan);
    if (tid == 0) write_column_plan(J, c, load_counter(&cnt[0]), st);
    __syncthreads();
    const uint32_t hist = st.hist;
    const int32_t vmin = st.vmin;
    if (!fixed) {
        for (int g = tid; g < G; g += nt) atomicExch(&cnt[g], 0u);
        __syncthreads();
        grouped([&](uint32_t g, V v, uint32_t) {
            if (!KT::finite(v) || v == V(0)) return;
            const uint32_t pos = off[g] + atomicAdd(&cnt[g], 1u);
            keys[pos] = hist ? (K)(uint32_t)((int)v - vmin) : KT::key(v);
        });
    } else if (hist) {
        // the slots hold keys; a histogram plan reads bins v - vmin
        const int w = tid / kWave, nw = nt / kWave;
        for (int g = w; g < G; g += nw) {
            const uint32_t o = off[g], m = min(load_counter(&cnt[g]), J.group_rows[g]);
            for (uint32_t i = (uint32_t)lane; i < m; i += kWave)
                keys[o + i] = (K)(uint32_t)((int)KT::value(keys[o + i]) - vmin);
        }
    }
}

// workgroup b takes queue item b: the split of jobs whose group counters do
// not fit LDS (kColBlock threads, several workgroups per CU)
template <typename V>
__global__ __launch_bounds__(kColBlock) void split_global_kernel(const JobDesc* jobs,
                                                                 const WorkItem* items,
                                                                 uint32_t n_items) {
    if (blockIdx.x < n_items) split_global_item<V, kColBlock>(jobs, items, blockIdx.x);
}

// Search of a sorted run a[base, base + n) in LDS with a fixed number of
// halvings (iters >= ceil(log2(n)), the same in every lane, so no lane waits
// on another's loop): the first index whose key is not below v (kUpper: the
// first above v).  Reads stay in a[0 .. last]; a halving of a run of one
// key or none reads a clamped index and leaves base alone.
template <bool kUpper, typename K>
__device__ __forceinline__ uint32_t bound_fixed(const HPDEX_LDS K* a, uint32_t base, uint32_t n,
                                                K v, uint32_t iters, uint32_t last) {
    for (uint32_t it = 0; it < iters; ++it) {
        const uint32_t half = n >> 1;
        const K x = a[min(base + half - 1u, last)];
        const bool right = kUpper ? !(v < x) : (x < v);
        base += right ? half : 0u;
        n -= half;
    }
    const K x = a[min(base, last)];
    const bool right = kUpper ? !(v < x) : (x < v);
    return base + ((n > 0u && right) ? 1u : 0u);
}

// Rank table of the reference slice seen by the target phase.
template <typename K>
struct RefView {
    const uint32_t* cum;     // histogram path: exclusive prefix of bin counts (LDS)
    const K* keys;           // sort path, reference beyond the on-chip sort: sorted keys in HBM
    const HPDEX_LDS K* lkeys;  // sort path: sorted keys in LDS
    const uint32_t* bidx;    // sort path: bucket -> first key index (LDS), nb + 1 entries
    // sort path, on chip: at the first of each run of equal keys, the run's
    // length (LDS), or null (the upper bound is searched instead)
    const HPDEX_LDS uint32_t* lrun;
    uint64_t kmin, kmax;
    uint32_t shift, na, hist, big, nb;
    uint32_t iters;          // halvings that cover the fullest bucket

    // lt = reference keys below v, eq = reference keys equal to v.  kWave:
    // every lane of the wave calls it together (the upper-bound search is
    // then skipped when no lane's key is present in the reference).
    template <bool kWave = false>
    __device__ inline void lookup(K v, uint32_t& lt, uint32_t& eq) const {
        if (hist) {
            const uint32_t a = cum[(uint32_t)v], b = cum[(uint32_t)v + 1];
            lt = a; eq = b - a;
            return;
        }
        if (big) {
            if (na == 0 || (uint64_t)v < kmin) { lt = 0; eq = 0; return; }
            if ((uint64_t)v > kmax) { lt = na; eq = 0; return; }
            const uint32_t bkt = (uint32_t)(((uint64_t)v - kmin) >> shift);
            const uint32_t lo = bidx[bkt], hi = bidx[bkt + 1];
            const uint32_t l = lower_bound_k<K>(gl(keys), lo, hi, v);
            const uint32_t u = upper_bound_k<K>(gl(keys), l, hi, v);
            lt = l; eq = u - l;
            return;
        }
        // below the smallest key: bucket 0, whose keys are all above v;
        // above the largest: the empty range at bidx[nb] = na
        const bool below = v < (K)kmin, above = v > (K)kmax;
        const uint32_t bkt = below ? 0u : above ? nb : (uint32_t)((K)(v - (K)kmin) >> shift);
        const uint32_t lo = bidx[bkt], hi = bidx[above ? bkt : bkt + 1u];
        const uint32_t last = na ? na - 1u : 0u;
        const uint32_t l = bound_fixed<false, K>(lkeys, lo, hi - lo, v, iters, last);
        const bool hit = l < hi && lkeys[min(l, last)] == v;
        if (lrun) {   // l is the first of the run of keys equal to v
            lt = l;
            eq = hit ? lrun[l] : 0u;
            return;
        }
        // the upper bound only where some lane's key is present
        uint32_t u = l;
        if (!kWave || __ballot(hit)) {
            const uint32_t uu = bound_fixed<true, K>(lkeys, l, hi - l, v, iters, last);
            u = hit ? uu : l;
        }
        lt = l; eq = u - l;
    }
};

// Halvings that take a run of up to maxb keys to one.
__device__ inline uint32_t halvings(uint32_t maxb) { return maxb <= 1u ? 0u : 32u - (uint32_t)__clz(maxb - 1u); }

// Sort reference keys rk[0..P) (P a power of two >= 64, padded with the
// sentinel) in LDS: registers for runs of up to kChunk, LDS merge stages above.
template <typename K>
__device__ __forceinline__ void sort_ref(K* rk, uint32_t P) {
    const int tid = tid_x(), nt = blockDim.x;
    const int wid = tid / kWave, nw = nt / kWave, lane = lane_id();
    if (P <= (uint32_t)kChunk) {
        if (wid == 0) {
            const int Kk = (int)(P / kWave);
#define HPDEX_SORT_SMALL(KK)                                                      \
    {                                                                              \
        K v[KK];                                                                   \
        for (int i = 0; i < KK; ++i) v[i] = rk[i * kWave + lane];                  \
        wave_sort<KK, K>(v);                                                       \
        for (int i = 0; i < KK; ++i) rk[i * kWave + lane] = v[i];                  \
    }
            if (Kk == 1) HPDEX_SORT_SMALL(1)
            else if (Kk == 2) HPDEX_SORT_SMALL(2)
            else if (Kk == 4) HPDEX_SORT_SMALL(4)
            else HPDEX_SORT_SMALL(8)
#undef HPDEX_SORT_SMALL
        }
        __syncthreads();
        return;
    }
    const uint32_t nchunk = P / kChunk;
    for (uint32_t ch = wid; ch < nchunk; ch += nw) {
        K v[kChunkK];
        K* b = rk + (size_t)ch * kChunk;
        for (int i = 0; i < kChunkK; ++i) v[i] = b[i * kWave + lane];
        wave_sort<kChunkK, K>(v, (int)(ch * kChunk));
        for (int i = 0; i < kChunkK; ++i) b[i * kWave + lane] = v[i];
    }
    __syncthreads();
    for (uint32_t k = 2 * kChunk; k <= P; k <<= 1) {
        for (uint32_t j = k >> 1; j >= (uint32_t)kChunk; j >>= 1) {
            for (uint32_t x = tid; x < P / 2; x += nt) {
                const uint32_t lo = (x / j) * 2 * j + (x % j);
                const uint32_t hi = lo + j;
                const bool asc = (lo & k) == 0;
                const K a = rk[lo], b = rk[hi];
                if (asc ? (b < a) : (a < b)) { rk[lo] = b; rk[hi] = a; }
            }
            __syncthreads();
        }
        for (uint32_t ch = wid; ch < nchunk; ch += nw) {
            K v[kChunkK];
            K* b = rk + (size_t)ch * kChunk;
            for (int i = 0; i < kChunkK; ++i) v[i] = b[i * kWave + lane];
            wave_merge_tail<kChunkK, K>(v, (int)k, (int)(ch * kChunk));
            for (int i = 0; i < kChunkK; ++i) b[i * kWave + lane] = v[i];
        }
        __syncthreads();
    }
}

// Sort the n keys a[0..n) in HBM (n above the cap keys an LDS sort takes):
// runs of cap keys sorted in LDS (rk, cap keys, cap a power of two), then
// merge passes of doubling width between a and b (n keys of spill space);
// the result ends in a.  Each thread merges one slice of a pair's output,
// found by a binary search along the merge path; equal keys take the left
// run's first.  Every barrier sits under workgroup-uniform control (n, cap).
template <typename K>
__device__ void sort_ref_global(K* a, K* b, uint32_t n, K* rk, uint32_t cap, K sentinel) {
    const uint32_t tid = (uint32_t)tid_x(), nt = blockDim.x;
    for (uint32_t q0 = 0; q0 < n; q0 += cap) {
        const uint32_t m = min(cap, n - q0);
        uint32_t P = 64;
        while (P < m) P <<= 1;
        for (uint32_t x = tid; x < P; x += nt) rk[x] = x < m ? a[q0 + x] : sentinel;
        __syncthreads();
        sort_ref<K>(rk, P);
        for (uint32_t x = tid; x < m; x += nt) a[q0 + x] = rk[x];
        __syncthreads();
    }
    K* src = a;
    K* dst = b;
    for (uint32_t w = cap; w < n; w <<= 1) {
        for (uint32_t i = 0; i < n; i += 2u * w) {
            const uint32_t nA = min(w, n - i);
            const uint32_t nB = n - i > w ? min(w, n - i - w) : 0u;
            const K* A = src + i;
            const K* B = A + nA;
            const uint32_t tot = nA + nB, per = (tot + nt - 1u) / nt;
            const uint32_t d0 = min(tid * per, tot), d1 = min(d0 + per, tot);
            if (d0 < d1) {
                // the first d0 outputs take ia keys of A: the least ia with
                // B[d0 - ia - 1] < A[ia]
                uint32_t lo = d0 > nB ? d0 - nB : 0u, hi = min(d0, nA);
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (!(B[d0 - mid - 1u] < A[mid])) lo = mid + 1u;
                    else hi = mid;
                }
                uint32_t ia = lo, ib = d0 - lo;
                K ka = ia < nA ? A[ia] : sentinel, kb = ib < nB ? B[ib] : sentinel;
                for (uint32_t d = d0; d < d1; ++d) {
                    const bool takeA = ib >= nB || (ia < nA && !(kb < ka));
                    dst[i + d] = takeA ? ka : kb;
                    if (takeA) { ++ia; ka = ia < nA ? A[ia] : sentinel; }
                    else { ++ib; kb = ib < nB ? B[ib] : sentinel; }
                }
            }
        }
        __syncthreads();
        K* t = src; src = dst; dst = t;
    }
    if (src != a) {
        for (uint32_t x = tid; x < n; x += nt) a[x] = src[x];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// reference: sort the reference slice once per gene and build its rank table
// (histogram prefix or bucket index) plus its own tie sum.
// ---------------------------------------------------------------------------
struct RefState {
    uint32_t scan[20];
    uint64_t red[16];
    // per-wave partial results of the on-chip branch, one slot per wave:
    // written by each wave's lane 0, read by every thread after a barrier
    uint64_t wlo[16], whi[16], wta[16];
    uint32_t wneg[16], wmaxb[16];
};

// Bucket sort of an on-chip reference: taken while its fullest bucket holds at
// most this many keys (each key is ranked by a pass over its bucket), else the
// bitonic sort
constexpr uint32_t kBucketSortMax = 64;

// Whether column c's reference tables come from the matrix's cache (the same
// reference rows give the same count and bucket count: checked, so a stale
// entry is never used).  Uniform over the workgroup.
__device__ __forceinline__ bool ref_cached(const JobDesc& J, int64_t c, uint32_t na, uint32_t nb) {
    if (!J.rc) return false;
    const RefCacheCol& e = J.rc[c];
    return uniform_u32(e.valid) == 1u && uniform_u32(e.na) == na && uniform_u32(e.nb) == nb;
}

template <typename V>
__device__ __forceinline__ void ref_item(const JobDesc* jobs, const WorkItem* items, uint32_t k) {
    using KT = KeyOf<V>;
    using K = typename KT::K;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ RefState st;
    int job;
    int64_t c;
    work_item(items, k, job, c);
    const JobDesc J = jobs[job];
    const int tid = tid_x(), nt = blockDim.x;
    const ColPlan pl = J.plan[c];
    K* base = (K*)J.keys + J.indptr[c];   // group 0's keys start the column
    uint32_t* tab = J.reftab + (size_t)c * kTabStride;
    // uniform decisions (kept in SGPRs so no barrier below sits under an exec mask)
    const uint32_t na = (uint32_t)__builtin_amdgcn_readfirstlane((int)pl.na);
    const uint32_t hist = (uint32_t)__builtin_amdgcn_readfirstlane((int)pl.hist);
    // the LDS/global sort cutover (a power of two, at most the on-chip capacity)
    const uint32_t cap = uniform_u32(J.ref_cap);
    // tables the matrix keeps for these reference rows: nothing to sort
    if (!hist && ref_cached(J, c, na, pl.nb)) {
        if (tid == 0) {
            const RefCacheCol& e = J.rc[c];
            ColPlan& q = J.plan[c];
            q.TA = e.TA;
            q.maxb = e.maxb;
            q.kmin = e.kmin;
            q.kmax = e.kmax;
            q.shift = e.shift;
            q.refneg = e.refneg;
        }
        return;
    }
    uint64_t ta = 0;
    uint32_t maxb = 0;   // keys in the fullest bucket of an on-chip reference
    // the reference's key range, bucket shift and values below zero (sort
    // path: read off the sorted keys; histogram branch: off its bins)
    uint64_t kmin = ~0ull, kmax = 0;
    uint32_t shift = 0, refneg = 0;
    // a reference slice beyond the on-chip sort: sorted in HBM with the
    // column's spill space, its bucket index built from the sorted keys there
    const int64_t spill_at =
        (!hist && na > cap) ? uniform_i64(J.spill_off ? J.spill_off[c] : -1) : -1;
    if (!hist && na > cap && spill_at < 0) {
        if (tid == 0) raise_status(J, kStatusCapacity, c);
    } else if (!hist && na > cap) {
        K* rk = (K*)smem;
        uint32_t* bidx = (uint32_t*)(rk + cap);
        sort_ref_global<K>(base, (K*)J.spill + spill_at, na, rk, cap, KT::kSentinel);
        kmin = (uint64_t)base[0];   // na > cap > 0
        kmax = (uint64_t)base[na - 1u];
        shift = ref_shift(kmin, kmax, na, pl.nb);
        for (uint32_t b = tid; b <= pl.nb; b += nt) bidx[b] = 0u;
        __syncthreads();
        uint64_t neg = 0;
        for (uint32_t x = tid; x < na; x += nt) {
            const K key = base[x];
            atomicAdd(&bidx[(uint32_t)(((uint64_t)key - kmin) >> shift)], 1u);
            neg += key < KT::kPosBit ? 1u : 0u;
        }
        __syncthreads();
        refneg = (uint32_t)block_sum_u64(neg, st.red);
        block_excl_scan(bidx, (int)pl.nb + 1, st.scan);
        for (uint32_t b = tid; b <= pl.nb; b += nt) tab[b] = bidx[b];
        RefView<K> rv;
        rv.cum = nullptr; rv.keys = base; rv.lkeys = nullptr; rv.bidx = bidx; rv.lrun = nullptr;
        rv.kmin = kmin; rv.kmax = kmax; rv.shift = shift; rv.na = na; rv.hist = 0;
        rv.big = 1; rv.nb = pl.nb; rv.iters = 0;
        for (uint32_t x = tid; x < na; x += nt) {
            if (x > 0 && base[x - 1] == base[x]) continue;   // not the first of its run
            uint32_t lt, eq;
            rv.lookup(base[x], lt, eq);
            ta += tie_f(eq);
        }
        ta = block_sum_u64(ta, st.red);
    } else if (hist) {
        uint32_t* cum = (uint32_t*)smem;
        for (uint32_t b = tid; b <= pl.nb; b += nt) cum[b] = 0u;
        __syncthreads();
        for (uint32_t x = tid; x < na; x += nt) atomicAdd(&cum[(uint32_t)base[x]], 1u);
        __syncthreads();
        for (uint32_t b = tid; b < pl.nb; b += nt) ta += tie_f(cum[b]);
        ta = block_sum_u64(ta, st.red);
        block_excl_scan(cum, (int)pl.nb + 1, st.scan);
        for (uint32_t b = tid; b <= pl.nb; b += nt) tab[b] = cum[b];
        // values below zero: the bins below bin -vmin
        if (pl.vmin < 0) refneg = cum[min((uint32_t)(-pl.vmin), pl.nb)];
    } else {
        // On chip: the key range (a block min and max), the bucket index
        // (counts, then their exclusive scan), then the sort -- by buckets
        // when the fullest bucket is small and a second key array fits:
        // each key goes to its bucket's range (the index as cursors), then
        // to its bucket's start plus the keys below it in the bucket plus
        // the equal keys before it there; the equal keys before it also give
        // its run's tie term, 3j(j - 1) for the j-th key of a run (summing
        // to t^3 - t over a run of t).  Otherwise the bitonic sort, and the
        // tie sum from lookups of the sorted keys.
        uint32_t P = 64;
        while (P < na) P <<= 1;
        K* rk = (K*)smem;
        uint32_t* bidx = (uint32_t*)(rk + P);
        const uint32_t idx_bytes = ((pl.nb + 1u) * 4u + 15u) & ~15u;
        K* tmp = (K*)((unsigned char*)bidx + idx_bytes);
        const bool room = (uint64_t)2u * P * sizeof(K) + idx_bytes <=
                          (uint64_t)ref_table_lds(sizeof(K) == 8);
        const uint32_t wid = (uint32_t)tid / (uint32_t)kWave, nw = (uint32_t)nt / (uint32_t)kWave;
        const bool lead = lane_id() == 0;
        K lo = KT::kSentinel, hi = K(0);
        uint32_t neg = 0;
        for (uint32_t x = tid; x < P; x += nt) {
            const K key = x < na ? base[x] : KT::kSentinel;
            rk[x] = key;
            if (x < na) {
                lo = key < lo ? key : lo;
                hi = key > hi ? key : hi;
                neg += key < KT::kPosBit ? 1u : 0u;
            }
        }
        for (uint32_t b = tid; b <= pl.nb; b += nt) bidx[b] = 0u;
        lo = wave_min_k<K>(lo);
        hi = wave_max_k<K>(hi);
        neg = (uint32_t)wave_sum_u64(neg);
        if (lead) { st.wlo[wid] = (uint64_t)lo; st.whi[wid] = (uint64_t)hi; st.wneg[wid] = neg; }
        __syncthreads();
        refneg = 0;
        for (uint32_t w = 0; w < nw; ++w) {
            kmin = st.wlo[w] < kmin ? st.wlo[w] : kmin;
            kmax = st.whi[w] > kmax ? st.whi[w] : kmax;
            refneg += st.wneg[w];
        }
        if (na == 0) { kmin = ~0ull; kmax = 0ull; }
        shift = ref_shift(kmin, kmax, na, pl.nb);
        for (uint32_t x = tid; x < na; x += nt)
            atomicAdd(&bidx[(uint32_t)(((uint64_t)rk[x] - kmin) >> shift)], 1u);
        __syncthreads();
        uint32_t m = 0;
        for (uint32_t b = tid; b < pl.nb; b += nt) m = max(m, bidx[b]);
        m = wave_max_u32(m);
        if (lead) st.wmaxb[wid] = m;
        block_excl_scan(bidx, (int)pl.nb + 1, st.scan);   // its barriers publish wmaxb
        for (uint32_t w = 0; w < nw; ++w) maxb = max(maxb, st.wmaxb[w]);
        maxb = uniform_u32(maxb);
        if (room && maxb <= kBucketSortMax) {
            for (uint32_t x = tid; x < na; x += nt) {
                const K key = rk[x];
                const uint32_t b = (uint32_t)(((uint64_t)key - kmin) >> shift);
                tmp[atomicAdd(&bidx[b], 1u)] = key;
            }
            __syncthreads();
            // bidx[b] now ends bucket b (the cursors moved one bucket on)
            for (uint32_t i = tid; i < na; i += nt) {
                const K key = tmp[i];
                const uint32_t b = (uint32_t)(((uint64_t)key - kmin) >> shift);
                const uint32_t s0 = b ? bidx[b - 1u] : 0u, s1 = bidx[b];
                uint32_t below = 0, before = 0;
                for (uint32_t j = s0; j < s1; ++j) {
                    const K y = tmp[j];
                    below += y < key ? 1u : 0u;
                    before += (y == key && j < i) ? 1u : 0u;
                }
                rk[s0 + below + before] = key;
                ta += 3ull * (uint64_t)(before + 1u) * (uint64_t)before;
            }
            ta = wave_sum_u64(ta);
            if (lead) st.wta[wid] = ta;
            __syncthreads();
            for (uint32_t x = tid; x < na; x += nt) base[x] = rk[x];
            for (uint32_t b = tid; b <= pl.nb; b += nt) tab[b] = b ? bidx[b - 1u] : 0u;
            ta = 0;
            for (uint32_t w = 0; w < nw; ++w) ta += st.wta[w];
        } else {
            sort_ref<K>(rk, P);
            for (uint32_t x = tid; x < na; x += nt) base[x] = rk[x];
            for (uint32_t b = tid; b <= pl.nb; b += nt) tab[b] = bidx[b];
            RefView<K> rv;
            rv.cum = nullptr; rv.keys = nullptr; rv.lkeys = lds(rk); rv.bidx = bidx;
            rv.lrun = nullptr;
            rv.kmin = kmin; rv.kmax = kmax; rv.shift = shift; rv.na = na; rv.hist = 0;
            rv.big = 0; rv.nb = pl.nb; rv.iters = halvings(maxb);
            for (uint32_t x = tid; x < na; x += nt) {
                uint32_t lt, eq;
                rv.lookup(rk[x], lt, eq);
                if (lt == x) ta += tie_f(eq);
            }
            ta = block_sum_u64(ta, st.red);
        }
    }
    if (tid == 0 && (hist || na <= cap || spill_at >= 0)) {
        ColPlan& q = J.plan[c];
        q.TA = ta;
        q.maxb = maxb;
        q.kmin = kmin;
        q.kmax = kmax;
        q.shift = shift;
        q.refneg = refneg;
    }
}

// workgroup b takes queue item b (the grid is sized for the queue's capacity)
template <typename V>
__global__ __launch_bounds__(kRefBlock) void ref_kernel(const JobDesc* jobs, const WorkItem* items,
                                                   uint32_t n_items) {
    if (blockIdx.x < n_items) ref_item<V>(jobs, items, blockIdx.x);
}

// Keep a finished job's reference tables in a matrix's cache: workgroup b
// copies column cols[b]'s sorted reference keys (group 0's keys at the start
// of the column's key range, sorted there by the reference kernel) and its
// bucket index to the offsets the host placed in rc[c], then records the
// fullest bucket and marks the entry valid.  The host filled koff, toff, na,
// nb and TA from the job's column plans.
template <typename K>
__global__ __launch_bounds__(kColBlock) void refcache_fill_kernel(const JobDesc* job,
                                                             const int32_t* cols, uint32_t n,
                                                             RefCacheCol* rc, K* rc_keys,
                                                             uint32_t* rc_tab) {
    __shared__ uint32_t maxb;
    if (blockIdx.x >= n) return;
    const uint32_t tid = (uint32_t)tid_x(), nt = blockDim.x;
    const int64_t c = cols[blockIdx.x];
    const RefCacheCol e = rc[c];
    const K* src = (const K*)job->keys + job->indptr[c];
    const uint32_t* tab = job->reftab + (size_t)c * kTabStride;
    for (uint32_t x = tid; x < e.na; x += nt) rc_keys[e.koff + x] = src[x];
    if (tid == 0) maxb = 0u;
    __syncthreads();
    uint32_t m = 0;
    for (uint32_t b = tid; b <= e.nb; b += nt) {
        const uint32_t lo = tab[b];
        rc_tab[e.toff + b] = lo;
        if (b < e.nb) m = max(m, tab[b + 1] - lo);
    }
    m = wave_max_u32(m);
    if (lane_id() == 0) atomicMax(&maxb, m);
    __syncthreads();
    if (tid == 0) {
        rc[c].maxb = maxb;
        rc[c].valid = 1u;
    }
}

// ---------------------------------------------------------------------------
// target: every target of a column against its rank table
// ---------------------------------------------------------------------------
template <typename K>
struct TargetCtx {
    RefView<K> ref;
    int32_t mode;
    int32_t vmin;          // histogram path: bin -> value offset
    uint32_t za;           // reference zero-block size
    uint64_t S;            // running 2*U_target (per lane)
    uint64_t TT;           // running tie delta (per lane)
};

// Is the zero block strictly below value key v?
template <typename V>
__device__ inline bool zero_below(const TargetCtx<typename KeyOf<V>::K>& t,
                                  typename KeyOf<V>::K v) {
    if (t.mode == kZeroMin) return true;
    if (t.mode == kZeroMax) return false;
    if (t.ref.hist) return (int32_t)v + t.vmin > 0;
    return v >= KeyOf<V>::kPosBit;
}

// Sort one run of up to 64*KK target keys in registers and accumulate its
// contribution.  seg: the target's keys in HBM; q: chunk number; earlier
// chunks (already sorted and written back) are searched for equal keys so
// that j counts the position inside the whole run of equal values.
template <typename V, int KK>
__device__ __forceinline__ void target_chunk(TargetCtx<typename KeyOf<V>::K>& t,
                                             typename KeyOf<V>::K* seg, uint32_t m, int q,
                                             bool store) {
    using K = typename KeyOf<V>::K;
    const int lane = lane_id();
    K* base = seg + (size_t)q * kChunk;
    K v[KK];
#pragma unroll
    for (int i = 0; i < KK; ++i) {
        const uint32_t e = (uint32_t)(i * kWave + lane);
        v[i] = e < m ? gl(base)[e] : KeyOf<V>::kSentinel;
    }
    wave_sort<KK, K>(v);
    if (store) {
#pragma unroll
        for (int i = 0; i < KK; ++i) {
            const uint32_t e = (uint32_t)(i * kWave + lane);
            if (e < m) gl(base)[e] = v[i];
        }
        __threadfence();
    }
    uint32_t carry = 0;
#pragma unroll
    for (int i = 0; i < KK; ++i) {
        const uint32_t e = (uint32_t)(i * kWave + lane);
        // previous element in sort order; lane 0 takes lane 63 of slot i-1
        const K last_prev = (i > 0) ? shfl_k<K>(v[i > 0 ? i - 1 : 0], kWave - 1) : v[i];
        K prev = shfl_up_k<K>(v[i], 1);
        if (lane == 0) prev = last_prev;
        const bool valid = e < m;
        const bool head = valid && (e == 0 || prev != v[i]);
        const uint64_t mask = __ballot(head);
        const uint64_t upto = (lane == kWave - 1) ? ~0ull : ((2ull << lane) - 1ull);
        const uint64_t below = mask & upto;
        const uint32_t h = below ? (uint32_t)(i * kWave + 63 - __clzll(below)) : carry;
        if (mask) carry = (uint32_t)(i * kWave + 63 - __clzll(mask));
        if (valid) {
            uint32_t lt, eq;
            t.ref.lookup(v[i], lt, eq);
            uint32_t prior = 0;
            for (int qq = 0; qq < q; ++qq) {
                const K* pc = seg + (size_t)qq * kChunk;
                const uint32_t l = lower_bound_k<K>(gl(pc), 0, kChunk, v[i]);
                const uint32_t u = upper_bound_k<K>(gl(pc), l, kChunk, v[i]);
                prior += u - l;
            }
            const uint64_t jj = (uint64_t)(e - h + 1) + prior;
            const uint64_t ltz = (uint64_t)lt + (zero_below<V>(t, v[i]) ? t.za : 0u);
            t.S += 2ull * ltz + eq;
            const uint64_t x = (uint64_t)eq + jj;
            t.TT += 3ull * x * (x - 1ull);
        }
    }
}

// Result of one target, the same in every lane of the wave.
struct TargetResult {
    double U1, tie;
    uint32_t n1, n2;
    int64_t idx;
};

template <typename V>
__device__ __forceinline__ TargetResult process_target(const JobDesc& J, int64_t c, uint32_t g,
                                                       const ColPlan& pl,
                                                       const RefView<typename KeyOf<V>::K>& ref) {
    using K = typename KeyOf<V>::K;
    const int64_t b = c * J.G;
    const uint32_t nzb = gl(J.cnt)[b + g];
    K* seg = (K*)J.keys + gl(J.indptr)[c] + gl(J.goff)[b + g];
    uint32_t za, zb, n1, n2;
    pair_sizes(J, c, g, za, zb, n1, n2);
    TargetCtx<K> t;
    t.ref = ref;
    t.mode = (int32_t)uniform_u32((uint32_t)J.mode);
    t.vmin = pl.vmin;
    t.za = za;
    t.S = 0;
    t.TT = 0;
    const uint32_t nch = (nzb + kChunk - 1) / kChunk;
    for (uint32_t q = 0; q < nch; ++q) {
        const uint32_t m = min((uint32_t)kChunk, nzb - q * kChunk);
        const bool store = nch > 1;
        if (m <= 64) target_chunk<V, 1>(t, seg, m, q, store);
        else if (m <= 128) target_chunk<V, 2>(t, seg, m, q, store);
        else if (m <= 256) target_chunk<V, 4>(t, seg, m, q, store);
        else target_chunk<V, 8>(t, seg, m, q, store);
    }
    // butterfly sums: every lane holds the totals
    const uint64_t S = wave_sum_u64(t.S);
    const uint64_t TT = wave_sum_u64(t.TT);
    uint32_t ltz;
    if (J.mode == kZeroMin) ltz = 0;
    else if (J.mode == kZeroMax) ltz = pl.na;
    else ltz = pl.refneg;
    const uint64_t S2 = S + (uint64_t)zb * (2ull * ltz + za);
    const uint64_t tie = TT + pl.TA + tie_f((uint64_t)za + zb);
    const uint64_t twoN = 2ull * (uint64_t)n1 * (uint64_t)n2;
    const uint64_t u1x2 = twoN >= S2 ? twoN - S2 : 0ull;
    TargetResult r;
    r.idx = c * J.T + (int64_t)(g - 1);
    r.U1 = (double)u1x2 * 0.5;
    r.tie = J.tie_corr ? (double)tie : 0.0;
    r.n1 = n1;
    r.n2 = n2;
    return r;
}

// process_target for a target of at most kHashTargetKeys 4-byte keys, without
// a sort: its keys in chunks of 64, one per lane, each ranked against the
// reference; a key's position among the target's equal keys comes from the
// wave's table htab (kBigTieSlots words, all empty on entry and on return):
// the key is placed by linear probing (compare and swap), then a counter
// beside it is incremented, the count before the increment being the
// position.  At most kHashTargetKeys keys in kBigTieSlots > kHashTargetKeys
// slots: every lane finds its key's slot or an empty one.
template <typename V>
__device__ __forceinline__ TargetResult process_target_hash(
    const JobDesc& J, int64_t c, uint32_t g, const ColPlan& pl,
    const RefView<typename KeyOf<V>::K>& ref, unsigned long long* htab) {
    using K = typename KeyOf<V>::K;
    static_assert(sizeof(K) == 4, "the tie table holds 32-bit keys");
    const uint32_t lane = (uint32_t)lane_id();
    const int64_t b = c * J.G;
    const uint32_t nzb = uniform_u32(gl(J.cnt)[b + g]);
    const K* seg = (const K*)J.keys + gl(J.indptr)[c] + gl(J.goff)[b + g];
    uint32_t za, zb, n1, n2;
    pair_sizes(J, c, g, za, zb, n1, n2);
    TargetCtx<K> t;
    t.ref = ref;
    t.mode = (int32_t)uniform_u32((uint32_t)J.mode);
    t.vmin = pl.vmin;
    t.za = za;
    uint64_t S = 0, TT = 0;
    constexpr uint32_t kChunks = kHashTargetKeys / (uint32_t)kWave;
#pragma unroll 1
    for (uint32_t q = 0; q < kChunks; ++q) {
        const uint32_t e = q * (uint32_t)kWave + lane;
        const bool valid = e < nzb;
        const K key = valid ? gl(seg)[e] : K(0);
        uint32_t lt, eq;
        t.ref.template lookup<true>(key, lt, eq);
        const uint64_t tag = (uint64_t)(uint32_t)key << 32;
        uint32_t sl = ((uint32_t)key * 0x9E3779B1u) >> 24;   // kBigTieSlots = 2^8
        bool placed = !valid;
        for (uint32_t probe = 0; probe < kBigTieSlots; ++probe) {
            if (__ballot(!placed) == 0ull) break;
            if (!placed) {
                const uint64_t old = atomicCAS(&htab[sl], (unsigned long long)kTieEmpty,
                                               (unsigned long long)tag);
                if (old == kTieEmpty || (old & ~kTieCountMask) == tag) placed = true;
                else sl = (sl + 1u) & (kBigTieSlots - 1u);
            }
        }
        if (valid) {
            const uint32_t jj = 1u + (uint32_t)(atomicAdd(&htab[sl], 1ull) & kTieCountMask);
            const uint64_t ltz = (uint64_t)lt + (zero_below<V>(t, key) ? za : 0u);
            S += 2ull * ltz + eq;
            const uint64_t x = (uint64_t)eq + jj;
            TT += 3ull * x * (x - 1ull);
        }
    }
    S = wave_sum_u64(S);
    TT = wave_sum_u64(TT);
    // the whole table empty again, after every lane's increments (the wave's
    // LDS operations complete in order)
    for (uint32_t i = lane; i < kBigTieSlots; i += (uint32_t)kWave) htab[i] = kTieEmpty;
    uint32_t ltz;
    if (J.mode == kZeroMin) ltz = 0;
    else if (J.mode == kZeroMax) ltz = pl.na;
    else ltz = pl.refneg;
    const uint64_t S2 = S + (uint64_t)zb * (2ull * ltz + za);
    const uint64_t tie = TT + pl.TA + tie_f((uint64_t)za + zb);
    const uint64_t twoN = 2ull * (uint64_t)n1 * (uint64_t)n2;
    const uint64_t u1x2 = twoN >= S2 ? twoN - S2 : 0ull;
    TargetResult r;
    r.idx = c * J.T + (int64_t)(g - 1);
    r.U1 = (double)u1x2 * 0.5;
    r.tie = J.tie_corr ? (double)tie : 0.0;
    r.n1 = n1;
    r.n2 = n2;
    return r;
}

// Outputs of up to 64 targets held one per lane: written, and their FP64
// asymptotic p computed, by the lanes in parallel.
__device__ __forceinline__ void flush_targets(const JobDesc& J, const TargetResult& r, int cnt) {
    if (lane_id() < cnt && r.idx >= 0) {
        gl(J.U1)[r.idx] = r.U1;
        if (J.tie) gl(J.tie)[r.idx] = r.tie;
        if (J.n1o) { gl(J.n1o)[r.idx] = (double)r.n1; gl(J.n2o)[r.idx] = (double)r.n2; }
        if (J.asym)
            gl(J.P)[r.idx] = p_asymptotic(r.U1, (double)r.n1, (double)r.n2, r.tie, J.cc, J.alt);
    }
}

// Targets g0 .. g0+ng-1 of column c, target g0+j held by lane j (4-byte keys).
// A window of consecutive targets whose keys total at most 64 is ranked in one
// pass of the wave: lane e finds the target of key e of the window (the last
// one starting at or before e in the window's packed order) and loads it from
// that target's keys; a key's position among the equal keys of its target
// comes from comparing it with the keys of the lanes before it in its
// segment, so nothing is sorted; inclusive scans of the per-key terms, read at
// each target's first and last position, give every target's sums.  Every
// cross-lane step but the key offset and those reads is a DPP move.  A target
// with more than 64 keys is left out (the window takes it as empty) and
// ranked apart.  Each lane returns its own target's result.
template <typename V>
__device__ __forceinline__ TargetResult target_windows(const JobDesc& J, int64_t c, uint32_t g0,
                                                       uint32_t ng, const ColPlan& pl,
                                                       const RefView<typename KeyOf<V>::K>& ref,
                                                       const typename KeyOf<V>::K* colkeys,
                                                       unsigned long long* htab) {
    using K = typename KeyOf<V>::K;
    static_assert(sizeof(K) == 4, "target windows compare 32-bit keys across lanes");
    const uint32_t lane = (uint32_t)lane_id();
    const int64_t b = c * J.G;
    const uint32_t gj = g0 + lane;
    const bool own = lane < ng;
    const uint32_t nzj = own ? gl(J.cnt)[b + gj] : 0u;
    const uint32_t off = own ? gl(J.goff)[b + gj] : 0u;
    // a target of more than 64 keys is ranked alone (process_target) after
    // the windows, which take it as empty
    const uint32_t sz = nzj <= (uint32_t)kWave ? nzj : 0u;
    uint32_t zaj = 0, zb = 0, n1 = 0, n2 = 0;
    if (own) pair_sizes(J, c, gj, zaj, zb, n1, n2);
    // the reference's zero block, the same for every target
    const uint32_t nza = uniform_u32(gl(J.cnt)[b]);
    const uint32_t gr0 = uniform_u32(gl(J.group_rows)[0]);
    const uint32_t za = uniform_u32((J.mode == kZeroNone) ? gl(J.zcnt)[b]
                                                          : (gr0 > nza ? gr0 - nza : 0u));
    const uint32_t inc = wave_incl_scan_dpp(sz, [](uint32_t a, uint32_t x) { return a + x; });
    // a window's per-key terms sum below 2^32 when 64 keys of 2*lt' + eq
    // (at most 2*(nza + za) + nza) and of 3x(x - 1) (x = eq + position, at
    // most nza + 64) do: 32-bit scans then
    const uint64_t xm = (uint64_t)nza + 64u;
    const bool narrow = 64ull * (2ull * ((uint64_t)nza + za) + nza) < (1ull << 32) &&
                        192ull * xm * xm < (1ull << 32);
    TargetCtx<K> t;
    t.ref = ref;
    t.mode = (int32_t)uniform_u32((uint32_t)J.mode);
    t.vmin = pl.vmin;
    t.za = za;
    uint64_t S = 0, TT = 0;
    bool big = false;
    TargetResult mine{};
    mine.idx = -1;   // nothing to write (a lane past the claim or a big target)
    // The window that starts at target j0: its end j1 (j1 == j0: target j0
    // alone has more than 64 keys), its m keys, and each lane's owner (the
    // target of its key), position of the owner's first key, and key.  The
    // loop plans window w + 1, issuing its key loads, before it ranks window
    // w: one window's loads are in flight while the one before is ranked.
    struct Win {
        uint32_t j0, j1, m, first, owner;
        K key;
    };
    // the start map: the 64 bytes after the wave's tie table (0xFF: no
    // target starts there)
    uint8_t* smap = (uint8_t*)(htab + kBigTieSlots);
    auto plan = [&](uint32_t j0) -> Win {
        Win w;
        w.j0 = j0; w.j1 = j0; w.m = 0; w.first = 0; w.owner = 0; w.key = K(0);
        if (j0 >= ng) return w;
        const uint32_t base = j0 ? readlane_u32(inc, j0 - 1u) : 0u;
        // at least target j0: no target takes more than 64 window keys
        const uint64_t fit = __ballot(lane >= j0 && own && inc - base <= (uint32_t)kWave);
        const uint32_t j1 = j0 + (uint32_t)__builtin_popcountll(fit);
        w.j1 = j1;
        w.m = readlane_u32(inc, j1 - 1u) - base;   // keys in the window
        // target j's keys take positions [excl_j, excl_j + sz_j) of the window
        const uint32_t excl = inc - sz - base;
        const bool starts = lane >= j0 && lane < j1 && sz > 0u;
        // the window's segment starts: each starting target writes its lane
        // at its first position in the wave's start map (64 bytes, 0xFF where
        // no target starts); the positions that hold a target, and the last
        // of them at or before a key's position, give the key's target.  The
        // wave's LDS operations complete in order (the ballots between them
        // order the lanes of a host model of the wave)
        if (starts) smap[excl] = (uint8_t)lane;
        (void)__ballot(starts);
        const uint64_t smask = __ballot(smap[lane] != 0xFFu);
        const uint64_t upto = (lane == (uint32_t)kWave - 1) ? ~0ull : ((2ull << lane) - 1ull);
        const uint64_t sm = smask & upto;
        w.first = sm ? (uint32_t)(63 - __clzll(sm)) : 0u;   // its first key
        w.owner = min((uint32_t)smap[w.first], (uint32_t)kWave - 1u);
        // the lane's key: position lane - first of the owner's keys (a
        // group's keys need not end where the next group's start)
        const uint32_t at = shfl_u32(off, (int)w.owner) + lane - w.first;
        if (starts) smap[excl] = 0xFFu;   // the map empty again
        w.key = lane < w.m ? gl(colkeys)[at] : K(0);
        return w;
    };
    Win w = plan(0u);
    {
        // the first window's keys in registers before the loop, so that the
        // loop waits for no key but the next window's, at its end
        K k0 = w.key;
        asm volatile("" : "+v"(k0));
        w.key = k0;
    }
    while (w.j0 < ng) {
        const Win nx = plan(w.j1);
        const uint32_t j0 = w.j0, j1 = w.j1;
        const uint32_t base = j0 ? readlane_u32(inc, j0 - 1u) : 0u;
        const bool valid = lane < w.m;
        const uint32_t excl = inc - sz - base;
        const bool starts = lane >= j0 && lane < j1 && sz > 0u;
        const uint32_t owner = w.owner;
        const K key = w.key;
        // every lane ranks (a lane past the window ranks key 0 and drops it)
        uint32_t lt, eq;
        t.ref.template lookup<true>(key, lt, eq);
        // the key's position among the equal keys of its target, plus one
        // (any order of equal keys gives the same sums): the (key, owner)
        // pair is placed in the wave's hash table (linear probing, compare
        // and swap), then a counter beside the pair is incremented; the
        // count before the increment is the position.  At most 64 pairs in
        // kBigTieSlots slots (a quarter full: short probe runs, and the loop
        // runs as long as the wave's longest): every lane finds its pair's
        // slot or an empty one within kBigTieSlots probes.
        const uint64_t tag = ((uint64_t)(uint32_t)key << 32) | ((uint64_t)owner << kTieCountBits);
        uint32_t slot = (((uint32_t)key * 0x9E3779B1u) ^ (owner * 0x85EBCA77u)) >> 24;
        bool placed = !valid;
        for (uint32_t probe = 0; probe < kBigTieSlots; ++probe) {
            if (__ballot(!placed) == 0ull) break;
            if (!placed) {
                const uint64_t old = atomicCAS(&htab[slot], (unsigned long long)kTieEmpty,
                                               (unsigned long long)tag);
                if (old == kTieEmpty || (old & ~kTieCountMask) == tag) placed = true;
                else slot = (slot + 1u) & (kBigTieSlots - 1u);
            }
        }
        uint32_t jj = 1u;
        if (valid) jj += (uint32_t)(atomicAdd(&htab[slot], 1ull) & kTieCountMask);
        uint64_t s_e = 0, t_e = 0;
        if (valid) {
            const uint64_t ltz = (uint64_t)lt + (zero_below<V>(t, key) ? za : 0u);
            s_e = 2ull * ltz + eq;
            const uint64_t x = (uint64_t)eq + jj;
            t_e = 3ull * x * (x - 1ull);
        }
        // inclusive scans of the terms; each target's scans at its last key
        // (an empty target ends where the one before it ends; none before
        // the window's first key), less those of the target before it, one
        // lane down (every lane takes part in the shuffles; the select comes
        // after)
        const uint32_t end = excl + sz;
        auto sums = [&](auto word_tag) {
            using U = typename decltype(word_tag)::type;
            const auto add = [](U a, U x) { return (U)(a + x); };
            const U ps = wave_incl_scan_dpp((U)s_e, add), pt = wave_incl_scan_dpp((U)t_e, add);
            U ps_at, pt_at, ps_prev, pt_prev;
            if constexpr (sizeof(U) == 8) {
                ps_at = shfl_u64(ps, (int)((end - 1u) & 63u));
                pt_at = shfl_u64(pt, (int)((end - 1u) & 63u));
            } else {
                ps_at = shfl_u32(ps, (int)((end - 1u) & 63u));
                pt_at = shfl_u32(pt, (int)((end - 1u) & 63u));
            }
            const U ps_end = end ? ps_at : U(0), pt_end = end ? pt_at : U(0);
            if constexpr (sizeof(U) == 8) {
                ps_prev = dpp_shr1_u64(ps_end);
                pt_prev = dpp_shr1_u64(pt_end);
            } else {
                ps_prev = wave_shr1_u32(ps_end);
                pt_prev = wave_shr1_u32(pt_end);
            }
            if (starts) {
                S = (uint64_t)(U)(ps_end - (lane == j0 ? U(0) : ps_prev));
                TT = (uint64_t)(U)(pt_end - (lane == j0 ? U(0) : pt_prev));
            }
        };
        if (narrow) sums(TypeTag<uint32_t>{});
        else sums(TypeTag<uint64_t>{});
        // the table empty again for the next window (the wave's LDS
        // operations complete in order: after every lane's increment)
        if (valid) htab[slot] = kTieEmpty;
        w = nx;
    }
    // targets of more than 64 keys: ranked apart (process_target_hash or
    // the sorted pass)
    big = own && nzj > (uint32_t)kWave;
    if (own && !big) {
        uint32_t ltz;
        if (J.mode == kZeroMin) ltz = 0;
        else if (J.mode == kZeroMax) ltz = pl.na;
        else ltz = pl.refneg;
        const uint64_t S2 = S + (uint64_t)zb * (2ull * ltz + za);
        const uint64_t tie = TT + pl.TA + tie_f((uint64_t)za + zb);
        const uint64_t twoN = 2ull * (uint64_t)n1 * (uint64_t)n2;
        const uint64_t u1x2 = twoN >= S2 ? twoN - S2 : 0ull;
        mine.idx = c * J.T + (int64_t)(gj - 1);
        mine.U1 = (double)u1x2 * 0.5;
        mine.tie = J.tie_corr ? (double)tie : 0.0;
        mine.n1 = n1;
        mine.n2 = n2;
    }
    return mine;
}

struct TargetState {
    uint32_t next_t;
    uint32_t next_h;
    uint32_t any_big;
};

// The target phase of one column.  kWindows (4-byte keys): a wave claims 64
// consecutive targets, one per lane, ranks those of at most 64 keys several
// at a time (target_windows) and those of 65 to kHashTargetKeys one at a time
// without a sort (process_target_hash); a column with larger targets is
// flagged for the sorted pass.  Otherwise (the sorted pass): a wave claims
// 64 targets and ranks each of them alone (process_target) -- every target
// for 8-byte keys, the targets of more than kHashTargetKeys keys of a flagged
// column for 4-byte keys.
template <typename V, bool kWindows>
__device__ __forceinline__ void target_item(const JobDesc* jobs, const WorkItem* items, uint32_t k) {
    using K = typename KeyOf<V>::K;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ TargetState st;
    int job;
    int64_t c;
    work_item(items, k, job, c);
    const JobDesc J = jobs[job];
    const int tid = tid_x(), nt = blockDim.x, lane = lane_id();
    const ColPlan pl = J.plan[c];
    // the sorted pass of 4-byte keys: only columns the window pass flagged
    if (!kWindows && sizeof(K) == 4 && uniform_u32(pl.bigt) == 0u) return;
    // uniform decisions (kept in SGPRs so the barrier below is not under an exec mask)
    const uint32_t na = (uint32_t)__builtin_amdgcn_readfirstlane((int)pl.na);
    const uint32_t hist = (uint32_t)__builtin_amdgcn_readfirstlane((int)pl.hist);
    // the LDS/global sort cutover (a power of two, at most the on-chip capacity)
    const uint32_t cap = uniform_u32(J.ref_cap);
    // the sorted reference and its table: the column's own (the reference
    // kernel sorted group 0's keys in place), or the matrix's cached ones
    const bool cached = !hist && ref_cached(J, c, na, pl.nb);
    const uint32_t* tab = cached ? J.rc_tab + J.rc[c].toff : J.reftab + (size_t)c * kTabStride;
    const K* rbase = cached ? (const K*)J.rc_keys + J.rc[c].koff
                            : (const K*)J.keys + gl(J.indptr)[c];
    // a reference beyond the on-chip sort: its sorted keys stay in HBM (read
    // through L2 by the lookups), its bucket index comes to LDS; with no spill
    // space the reference kernel raised the status
    const bool big = !hist && na > cap;
    if (big && !cached && uniform_i64(J.spill_off ? gl(J.spill_off)[c] : -1) < 0) return;
    RefView<K> rv;
    rv.kmin = pl.kmin; rv.kmax = pl.kmax; rv.shift = pl.shift; rv.na = na; rv.hist = hist;
    rv.big = big ? 1u : 0u; rv.nb = pl.nb; rv.iters = 0; rv.lkeys = nullptr; rv.lrun = nullptr;
    // an on-chip reference filling at most half the key space, whose runs of
    // equal keys (each inside one bucket) hold at most kBucketSortMax keys:
    // the run lengths go to the free half, so a present key's equal count is
    // one read instead of a second search
    uint32_t* runs = nullptr;
    K* rkeys = nullptr;
    if (hist) {
        uint32_t* cum = (uint32_t*)smem;
        for (uint32_t b = tid; b <= pl.nb; b += nt) cum[b] = gl(tab)[b];
        rv.cum = cum; rv.keys = nullptr; rv.bidx = nullptr;
    } else if (big) {
        uint32_t* bidx = (uint32_t*)((K*)smem + cap);
        for (uint32_t b = tid; b <= pl.nb; b += nt) bidx[b] = gl(tab)[b];
        rv.cum = nullptr; rv.keys = rbase; rv.bidx = bidx;
    } else {
        K* rk = (K*)smem;
        uint32_t* bidx = (uint32_t*)(rk + cap);
        const K* base = rbase;
        for (uint32_t x = tid; x < na; x += nt) rk[x] = gl(base)[x];
        for (uint32_t b = tid; b <= pl.nb; b += nt) bidx[b] = gl(tab)[b];
        rv.cum = nullptr; rv.keys = nullptr; rv.lkeys = lds(rk); rv.bidx = bidx;
        rv.iters = halvings(uniform_u32(pl.maxb));
        if (2u * na <= cap && uniform_u32(pl.maxb) <= kBucketSortMax) {
            runs = (uint32_t*)(rk + na);
            rkeys = rk;
            rv.lrun = lds(runs);
        }
    }
    // after the barrier that publishes the keys: the run lengths, then a
    // barrier of their own (runs is uniform)
    auto fill_runs = [&]() {
        if (!runs) return;
        for (uint32_t x = tid; x < na; x += nt) {
            if (x > 0 && rkeys[x - 1u] == rkeys[x]) continue;   // not the first of its run
            uint32_t e = x + 1u;
            while (e < na && rkeys[e] == rkeys[x]) ++e;
            runs[x] = e - x;
        }
        __syncthreads();
    };
    if (tid == 0) { st.next_t = 1u; st.next_h = 1u; st.any_big = 0u; }
    if constexpr (kWindows) {
        static_assert(sizeof(K) == 4, "target windows take 4-byte keys");
        // each wave's tie-position table (target_windows, process_target_hash)
        // and start map, all slots empty (all-ones words)
        unsigned long long* htab = (unsigned long long*)(smem + ref_table_lds(false)) +
                                   (size_t)(tid / kWave) * kWaveTableWords;
        for (uint32_t i = (uint32_t)lane; i < kWaveTableWords; i += (uint32_t)kWave)
            htab[i] = kTieEmpty;
        __syncthreads();
        fill_runs();
        const K* colkeys = (const K*)J.keys + gl(J.indptr)[c];
        bool any_big = false;
        for (;;) {
            uint32_t g0 = 0;
            if (lane == 0) g0 = atomicAdd(&st.next_t, (uint32_t)kWave);
            g0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)shfl_u32(g0, 0));
            if (g0 >= (uint32_t)J.G) break;
            const uint32_t ng = min((uint32_t)kWave, (uint32_t)J.G - g0);
            flush_targets(J, target_windows<V>(J, c, g0, ng, pl, rv, colkeys, htab), (int)ng);
        }
        // then the targets of 65 to kHashTargetKeys keys, ranked without a
        // sort and each written by its lane (a second claim loop: nothing of
        // the windows is live here); larger ones are flagged for the sorted
        // pass
        for (;;) {
            uint32_t g0 = 0;
            if (lane == 0) g0 = atomicAdd(&st.next_h, (uint32_t)kWave);
            g0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)shfl_u32(g0, 0));
            if (g0 >= (uint32_t)J.G) break;
            const uint32_t ng = min((uint32_t)kWave, (uint32_t)J.G - g0);
            const uint32_t nzj =
                (uint32_t)lane < ng ? gl(J.cnt)[c * J.G + g0 + (uint32_t)lane] : 0u;
            for (uint64_t m = __ballot(nzj > (uint32_t)kWave && nzj <= kHashTargetKeys); m;
                 m &= m - 1ull) {
                const uint32_t j = (uint32_t)__builtin_ctzll(m);
                TargetResult h = process_target_hash<V>(J, c, g0 + j, pl, rv, htab);
                if ((uint32_t)lane != j) h.idx = -1;
                flush_targets(J, h, (int)ng);
            }
            any_big = any_big || __ballot(nzj > kHashTargetKeys) != 0ull;
        }
        if (any_big && lane == 0) st.any_big = 1u;
        __syncthreads();
        if (tid == 0 && st.any_big) J.plan[c].bigt = 1u;
    } else {
        __syncthreads();
        fill_runs();
        // lane j of the wave keeps target g0 + j's result
        for (;;) {
            uint32_t g0 = 0;
            if (lane == 0) g0 = atomicAdd(&st.next_t, (uint32_t)kWave);
            g0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)shfl_u32(g0, 0));
            if (g0 >= (uint32_t)J.G) break;
            const uint32_t ng = min((uint32_t)kWave, (uint32_t)J.G - g0);
            const bool own = (uint32_t)lane < ng;
            const uint32_t nzj = own ? gl(J.cnt)[c * J.G + g0 + (uint32_t)lane] : 0u;
            TargetResult mine{};
            mine.idx = -1;
            for (uint64_t m = __ballot(own && (sizeof(K) == 8 || nzj > kHashTargetKeys)); m;
                 m &= m - 1ull) {
                const uint32_t j = (uint32_t)__builtin_ctzll(m);
                const TargetResult r = process_target<V>(J, c, g0 + j, pl, rv);
                if ((uint32_t)lane == j) mine = r;
            }
            flush_targets(J, mine, (int)ng);
        }
    }
}

// The window pass of 4-byte keys: workgroup b takes queue item b (the grid
// is sized for the queue's capacity).  512 threads at four waves per SIMD
// (at most 128 VGPRs): the reference table and eight tie tables (65 KB of
// LDS) fit two workgroups per CU, 16 waves to hide the latency of the
// LDS searches and key loads.
template <typename V>
__global__ __launch_bounds__(kTargetBlock) __attribute__((amdgpu_waves_per_eu(4)))
void target_window_kernel(const JobDesc* jobs, const WorkItem* items, uint32_t n_items) {
    if (blockIdx.x < n_items) target_item<V, true>(jobs, items, blockIdx.x);
}

// The sorted pass (every target of 8-byte keys; the targets of more than
// kHashTargetKeys keys of 4-byte keys): workgroup b takes queue item b.  At most 168 VGPRs,
// so that three waves share a SIMD: the table LDS (49 KB) fits three
// workgroups per CU.
template <typename V>
__global__ __launch_bounds__(kColBlock) __attribute__((amdgpu_waves_per_eu(3)))
void target_kernel(const JobDesc* jobs, const WorkItem* items, uint32_t n_items) {
    if (blockIdx.x < n_items) target_item<V, false>(jobs, items, blockIdx.x);
}

// ---------------------------------------------------------------------------
// count: the whole pass of an integer-valued column in one workgroup (the
// UMI-count case of mannwhitneyu.cpp:903-1043).
//
// The resident matrix carries each column's value range [vmin, vmax] (stored
// nonzero finite values), so histogram bins b = value - vmin are known before
// the column is read.  One read of the column from HBM then does everything
// with one LDS atomic per stored entry into a table of counter rows: row 0
// counts the reference's values per bin, row r the values of target t0 + r,
// and an extra column NB of every row counts the group's explicit zeros.
// Afterwards, with lt(b)/eq(b) from the reference row's exclusive prefix and
// b_gv from the counters, each target's sums follow per bin:
//
//   2*U_target(g) += b_gv * (2*lt(v) + eq(v))      (zero block added below)
//   tie(g)        += f(eq(v) + b_gv) - f(eq(v))      f(t) = t^3 - t
//
// so no entry needs to know its position inside a run of equal values.  When
// the rows of every target do not fit the workgroup's LDS, the host splits
// the targets into slices of S (CountItem), one workgroup each: every slice
// reads the whole column and counts the reference row and its own targets,
// so the slices of a column run side by side rather than one after another.
// Columns whose slices would be too many take the split/reference/target
// kernels instead (the host routes them); a column holding a value outside
// its recorded range (the matrix's metadata no longer matches its values) is
// an error of the job (kStatusRange).
// ---------------------------------------------------------------------------
struct CountState {
    uint32_t bad;
    uint32_t pad[3];
    uint32_t scan[20];
    uint64_t red[16];
};

// Counters of (row, bin) pairs: 16-bit halves of LDS words when no count can
// reach 2^16 (a column of fewer than 2^16 entries), else whole words.  Rows
// are `rs` counters apart (count_row_stride: an odd number of words), so the
// per-target sums (one thread per row) read without bank conflicts.
struct PairCounters {
    uint32_t* w;
    uint32_t rs;
    bool wide;
    __device__ __forceinline__ uint32_t get(uint32_t row, uint32_t d) const {
        const uint32_t q = row * rs + d;
        return wide ? w[q] : ((w[q >> 1] >> ((q & 1u) << 4)) & 0xFFFFu);
    }
};

// One (column, target slice) item of a job: the column's extents and value
// range in `item_in`, the job's fields read with scalar loads into SGPRs
// (J is a kernel argument, or a job of the batch's descriptor array at a
// workgroup-uniform index).
template <typename V, int NT>
__device__ __forceinline__ void count_item(const JobDesc& J, const CountItem& item_in,
                                           uint32_t lds_bytes, CountState& st) {
    using KT = KeyOf<V>;
    extern __shared__ __align__(16) unsigned char smem[];
    // every value that steers control flow around a barrier is read into an
    // SGPR (a branch on a VGPR would put the barriers under exec masks)
    const CountItem item = item_in;
    const int64_t c = (int64_t)uniform_u32((uint32_t)item.col);
    const uint32_t t0 = uniform_u32((uint32_t)item.t0);
    const uint32_t S = uniform_u32((uint32_t)item.S);
    const int64_t p0 = uniform_i64(item.p0);
    const uint32_t n = uniform_u32(item.n);
    const int32_t vmin = (int32_t)uniform_u32((uint32_t)item.vmin);
    const uint32_t NB = uniform_u32(item.nb);   // histogram bins
    const int tid = tid_x(), nt = NT;
    const uint32_t G = uniform_u32((uint32_t)J.G), nT = G - 1u;
    const uint32_t R = uniform_u32((uint32_t)J.R);
    const uint32_t Ts = (uint32_t)__builtin_amdgcn_readfirstlane((int)min(S, nT - t0));
    const uint32_t rows = Ts + 1u;   // reference + this slice's targets
    const bool sliced = Ts != nT;
    const V* vals = (const V*)J.vals;
    const bool wide = n >= 65536u;

    // the same test on every slice of the column (rows of S + 1), so that they
    // all agree whether the column is theirs
    // (the host routed the column with the same tuning and metadata: a
    // column failing here would be a host/device disagreement, reported as
    // an error rather than counted wrongly)
    const int unfit = __builtin_amdgcn_readfirstlane(
        ((int64_t)NB > (int64_t)uniform_u32((uint32_t)J.hist_max) ||
         count_lds_need(S + 1u, NB, wide) > (uint64_t)lds_bytes) ? 1 : 0);
    if (unfit) {
        if (t0 == 0u && tid == 0) raise_status(J, kStatusRange, c);
        return;
    }

    // group sizes of the reference and this slice's targets, fetched now and
    // kept in registers through the read pass (row x: group 0 or t0 + x)
    const uint32_t* grows = J.group_rows;
    const uint32_t gpre0 = (uint32_t)tid < rows ? grows[tid == 0 ? 0u : t0 + (uint32_t)tid] : 0u;
    const uint32_t gpre1 = (uint32_t)tid + (uint32_t)nt < rows ? grows[t0 + (uint32_t)(tid + nt)] : 0u;

    // LDS: dummy[64] | acc[2*rows] u64 | cez, cnz, grs [rows] | bins[NB] |
    // counters[rows x rs]
    uint32_t* lds_words = (uint32_t*)smem;
    uint64_t* acc = (uint64_t*)(lds_words + kCountDummyWords);   // [2x] 2*U_target part,
                                                                  // [2x+1] tie part
    uint32_t* cez = (uint32_t*)(acc + 2u * rows);   // explicit zeros per row
    uint32_t* cnz = cez + rows;                     // nonzero values per row
    uint32_t* grs = cnz + rows;                     // group size per row
    uint2* bins = (uint2*)(cez + count_row_words(rows));   // per bin: (lt', eq), below
    const uint32_t rs = (uint32_t)__builtin_amdgcn_readfirstlane((int)count_row_stride(NB + 1u, wide));
    PairCounters pc;
    pc.w = (uint32_t*)(bins + NB);
    pc.rs = rs;
    pc.wide = wide;
    const uint32_t cell_words = count_cell_words(rows, rs, wide);
    const uint32_t cell_off = (uint32_t)(pc.w - lds_words);

    // column loads (buffer loads, described with the read pass below); the
    // first batch is issued now so that it overlaps the set-up of LDS.  A
    // packed column (kCountPacked) is read as one word per entry, (row << vb)
    // | code, from the matrix's packed entries instead of a row and a value.
    const bool packed = uniform_u32(item.flags & kCountPacked) != 0u;
    const uint32_t pvb = uniform_u32(J.pack_vb);
    const uint32_t pmask = (1u << pvb) - 1u;
    const int32_t* col_rows = J.rows + p0;
    const V* col_vals = vals + p0;
    const uint32_t* col_packed = J.packed + p0;
    const __amdgpu_buffer_rsrc_t gid_rsrc = __builtin_amdgcn_make_buffer_rsrc(
        (void*)J.gidp, (short)0, uniform_u32(J.gid_words) * 4u, kBufferWord3);
    const uint32_t lbits = uniform_u32(J.gid_bits);
    const uint32_t lmask = (1u << lbits) - 1u;
    const uint32_t step = (uint32_t)NT * kUnroll;
    bool bad = false;
    auto load = [&](uint32_t ib, uint32_t (&r)[kUnroll], V (&v)[kUnroll]) {
        // resources from the batch's first entry to the column's end: lanes
        // past the end read 0 without a memory access
        const __amdgpu_buffer_rsrc_t rows_rsrc = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(col_rows + ib), (short)0, (n - ib) * 4u, kBufferWord3);
        const __amdgpu_buffer_rsrc_t vals_rsrc = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(col_vals + ib), (short)0, (n - ib) * (uint32_t)sizeof(V), kBufferWord3);
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            // per lane: entry tid + u*NT of the batch (loop-invariant)
            const uint32_t e = (uint32_t)tid + (uint32_t)(u * NT);
            r[u] = buffer_load_u32(rows_rsrc, e * 4u, 0u);
            v[u] = buffer_load_value<V>(vals_rsrc, e * (uint32_t)sizeof(V), 0u);
        }
    };
    auto load_packed = [&](uint32_t ib, uint32_t (&r)[kUnroll]) {
        const __amdgpu_buffer_rsrc_t pk_rsrc = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(col_packed + ib), (short)0, (n - ib) * 4u, kBufferWord3);
#pragma unroll
        for (int u = 0; u < kUnroll; ++u)
            r[u] = buffer_load_u32(pk_rsrc, ((uint32_t)tid + (uint32_t)(u * NT)) * 4u, 0u);
    };
    // kClamp: rows may lie past R (unpacked rows of -1, masked lanes); a
    // packed word of a whole batch holds a row of at most R.  kR24: R < 2^24,
    // so the bit position takes the full-rate 24-bit multiply.
    auto gather = [&](const uint32_t (&r)[kUnroll], uint32_t (&g)[kUnroll], auto packed_tag,
                      auto clamp_tag, auto r24_tag) {
        constexpr bool kPk = decltype(packed_tag)::value;
        constexpr bool kClamp = decltype(clamp_tag)::value;
        constexpr bool kR24 = decltype(r24_tag)::value;
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            // the label's bit position; 8 bytes from its word hold all of it
            uint32_t row = kPk ? r[u] >> pvb : r[u];
            if (kClamp) row = min(row, R);
            const uint32_t pos = kR24 ? __umul24(row, lbits) : row * lbits;
            const uint64_t w2 = buffer_load_u64(gid_rsrc, (pos >> 5) * 4u);
            g[u] = (uint32_t)(w2 >> (pos & 31u)) & lmask;
        }
    };
    uint32_t r0[kUnroll];
    V v0[kUnroll];
    if (packed) load_packed(0u, r0);
    else load(0u, r0, v0);

    for (uint32_t x = tid; x < cell_words; x += nt) pc.w[x] = 0u;
    if (tid == 0) st.bad = 0u;
    __syncthreads();

    // the one read of the column from HBM.  Lane i of a wave takes entries i,
    // i+64, ... so a wave's 64 entries are adjacent: their row indices are
    // sorted, and the group labels they gather (bit-packed at ceil(log2(G+1))
    // bits, so a column's gathers touch as few cache lines as possible; rows
    // out of range read the sentinel at index R) fall into a few lines.
    // Batches of kUnroll rows and values per lane are software-pipelined: the
    // next batch's loads are issued after this batch's label gathers and
    // before its atomics, so they are in flight while this batch is counted
    // (loads complete in order: waiting for the gathers leaves them
    // outstanding).  Each entry costs one LDS atomic, its word and increment
    // computed without branches; entries of no row of this workgroup hit the
    // lane's dummy word.
    //
    // The loads are buffer loads: each batch's start and the bytes left to
    // the column's end in a scalar resource, the lane's offset in VGPRs fixed
    // for the whole loop.  The hardware range check turns the lanes past the
    // column's end into zeros that cost no memory traffic (they are masked in
    // the last batch), so no load leaves the column.  Columns are shorter
    // than 2^28 entries, so the byte sizes fit 32 bits.
    auto count = [&](const uint32_t (&g)[kUnroll], const V (&v)[kUnroll],
                     const uint32_t (&r)[kUnroll], auto wide_tag, auto sliced_tag,
                     auto packed_tag) {
        constexpr bool kWide = decltype(wide_tag)::value;
        constexpr bool kSliced = decltype(sliced_tag)::value;
        constexpr bool kPk = decltype(packed_tag)::value;
        const uint32_t dummy = (uint32_t)lane_id();
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const uint32_t gg = g[u];
            uint32_t b;
            bool zero, inb;
            if (kPk) {
                // the code: a value's bin, NB for an explicit zero (the zero
                // column), or the skip code of a non-finite value -- so the
                // counter is row * rs + code for whatever is counted, and no
                // code lies outside the bins
                b = r[u] & pmask;
                zero = false;
                inb = b <= NB;
            } else {
                const V vv = v[u];
                // |v| <= 32768 covers every integer the column metadata
                // admits ([-32768, 32767]); a finite nonzero value outside
                // the recorded range means the metadata no longer matches
                // (checked for every entry, whatever its group, so all
                // slices agree)
                const bool near = KT::abs_le(vv, V(32768));
                zero = vv == V(0);
                b = (uint32_t)((int)(near ? vv : V(0)) - vmin);
                inb = near && b < NB;
                bad |= !zero && !inb && KT::finite(vv);
            }
            // the entry's row: reference 0, target t0 + k -> k
            uint32_t row;
            bool mine;
            if (kSliced) {
                const uint32_t rr = gg - t0;
                row = gg == 0u ? 0u : rr;
                mine = gg == 0u || rr - 1u < Ts;
            } else {
                row = gg;
                mine = gg < G;
            }
            // row < 2^16 whenever the entry is counted (a garbage row only
            // feeds a word that is not used)
            const uint32_t q = __umul24(row, rs) + (kPk ? b : (zero ? NB : b));
            const uint32_t word = (mine && (zero || inb)) ? cell_off + (kWide ? q : (q >> 1)) : dummy;
            atomicAdd(&lds_words[word], kWide ? 1u : 1u + (q & 1u) * 65535u);
        }
    };
    auto read_all = [&](auto wide_tag, auto sliced_tag, auto packed_tag, auto r24_tag) {
        constexpr bool kPk = decltype(packed_tag)::value;
        // two register sets: each batch is counted from the set it was loaded
        // into while the next one loads into the other (no copies)
        uint32_t rA[kUnroll], rB[kUnroll];
        V vA[kUnroll], vB[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) { rA[u] = r0[u]; vA[u] = kPk ? V(0) : v0[u]; }
        // a whole batch with another after it
        auto whole = [&](const uint32_t (&rc)[kUnroll], const V (&vc)[kUnroll],
                         uint32_t (&rn)[kUnroll], V (&vn)[kUnroll], uint32_t ibn) {
            uint32_t g[kUnroll];
            gather(rc, g, packed_tag, std::integral_constant<bool, !kPk>{}, r24_tag);
            // keep the gathers ahead of the next batch's loads (the counter
            // that waits for the gathers would otherwise wait for those too)
            __builtin_amdgcn_sched_barrier(0);
            if (kPk) load_packed(ibn, rn);
            else load(ibn, rn, vn);
            __builtin_amdgcn_sched_barrier(0);
            count(g, vc, rc, wide_tag, sliced_tag, packed_tag);
        };
        // the last batch: entries past the column's end masked (a masked
        // packed word reads as row R with the skip code)
        auto last = [&](uint32_t (&rc)[kUnroll], V (&vc)[kUnroll], uint32_t ib) {
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                const bool in = ib + (uint32_t)(u * NT + tid) < n;
                rc[u] = in ? rc[u] : 0xFFFFFFFFu;
                vc[u] = in ? vc[u] : V(0);
            }
            uint32_t g[kUnroll];
            gather(rc, g, packed_tag, std::true_type{}, r24_tag);
            count(g, vc, rc, wide_tag, sliced_tag, packed_tag);
        };
        uint32_t ib = 0;
        for (;;) {
            if (ib + step >= n) {
                if (ib < n) last(rA, vA, ib);
                break;
            }
            whole(rA, vA, rB, vB, ib + step);
            ib += step;
            if (ib + step >= n) {
                last(rB, vB, ib);
                break;
            }
            whole(rB, vB, rA, vA, ib + step);
            ib += step;
        }
    };
    auto read_as = [&](auto packed_tag, auto r24_tag) {
        if (sliced) {
            if (wide) read_all(std::true_type{}, std::true_type{}, packed_tag, r24_tag);
            else read_all(std::false_type{}, std::true_type{}, packed_tag, r24_tag);
        } else {
            if (wide) read_all(std::true_type{}, std::false_type{}, packed_tag, r24_tag);
            else read_all(std::false_type{}, std::false_type{}, packed_tag, r24_tag);
        }
    };
    if (R < (1u << 24)) {
        if (packed) read_as(std::true_type{}, std::true_type{});
        else read_as(std::false_type{}, std::true_type{});
    } else {
        if (packed) read_as(std::true_type{}, std::false_type{});
        else read_as(std::false_type{}, std::false_type{});
    }
    if (__any(bad ? 1 : 0) && lane_id() == 0) atomicOr(&st.bad, 1u);
    if ((uint32_t)tid < rows) grs[tid] = gpre0;
    if ((uint32_t)tid + (uint32_t)nt < rows) grs[tid + nt] = gpre1;
    for (uint32_t x = (uint32_t)tid + 2u * (uint32_t)nt; x < rows; x += nt) grs[x] = grows[t0 + x];
    __syncthreads();
    // a value outside the recorded range: the matrix's metadata no longer
    // matches its values -- an error of the job (raised once, by the first slice)
    if (__builtin_amdgcn_readfirstlane((int)st.bad)) {
        if (t0 == 0u && tid == 0) raise_status(J, kStatusRange, c);
        return;
    }

    // one pass over the reference row: each thread takes a run of `chunk`
    // bins, the waves' inclusive scans and a scan of the wave totals give
    // every bin the number of reference values below it; with them the tie
    // sum of the reference values and the explicit zeros of every row
    const int32_t mode = (int32_t)uniform_u32((uint32_t)J.mode);
    const uint32_t chunk = (NB + (uint32_t)NT - 1u) / (uint32_t)NT;
    const uint32_t b0 = min((uint32_t)tid * chunk, NB), b1 = min(b0 + chunk, NB);
    uint32_t sh = 0;
    uint64_t stie = 0;
    for (uint32_t b = b0; b < b1; ++b) {
        const uint32_t h = pc.get(0u, b);
        sh += h;
        stie += tie_f(h);
    }
    for (uint32_t x = tid; x < rows; x += nt) cez[x] = pc.get(x, NB);
    const uint32_t incl = wave_incl_scan_u32(sh);
    const uint64_t wtie = wave_sum_u64(stie);
    const int w = tid / kWave;
    if (lane_id() == kWave - 1) st.scan[w] = incl;
    if (lane_id() == 0) st.red[w] = wtie;
    __syncthreads();
    uint32_t below = 0, na = 0;   // reference values in earlier waves' bins; in all bins
    uint64_t ta = 0;              // reference tie sum over nonzero values
#pragma unroll
    for (int i = 0; i < NT / kWave; ++i) {
        const uint32_t t = st.scan[i];
        below += i < w ? t : 0u;
        na += t;
        ta += st.red[i];
    }
    const uint32_t gr0 = grs[0];
    const uint32_t za = (mode == kZeroNone) ? cez[0] : (gr0 > na ? gr0 - na : 0u);
    // bins[b] = (2*lt' + eq, eq): lt' = reference values below b, plus the
    // zero block when it lies below b; eq = reference values in b.  lt' + eq
    // counts distinct reference rows, at most R < 2^31, so 2*lt' + eq fits
    // 32 bits.
    {
        uint32_t lt = below + incl - sh;
        for (uint32_t b = b0; b < b1; ++b) {
            const uint32_t h = pc.get(0u, b);
            // the first bin of a positive value (a packed column's bins are
            // ranks: vmin holds its distinct negative values)
            const bool pos = packed ? b >= (uint32_t)vmin : (int32_t)b + vmin > 0;
            const bool zbelow = mode == kZeroMin || (mode != kZeroMax && pos);
            bins[b] = make_uint2(2u * (lt + (zbelow ? za : 0u)) + h, h);
            // where 0 would sit: the reference values below it
            if (packed ? b == (uint32_t)vmin : (int32_t)b + vmin == 0) st.scan[16] = lt;
            lt += h;
        }
    }
    __syncthreads();

    // per-target sums over the bins, one thread per target row, no branches
    // per bin (a count of 0 adds 0 to each sum).  With c the row's count in
    // bin b, a = 2*lt' + eq and t = eq + c:
    //   2*U part: sum_b c * a
    //   tie part: sum_b f(t) - f(eq) = sum_b t^3 - (na + k) - ta
    // for f(t) = t^3 - t, since sum_b t = na + k (k = the row's count) and
    // sum_b f(eq) = ta, the reference's own tie sum over every bin; exact
    // modulo 2^64 like any other order of the same integer additions.  A
    // column of fewer than 2^16 entries has t < 2^16, so t^2 takes the
    // full-rate 24-bit multiply.
    for (uint32_t x = 1u + tid; x < rows; x += nt) {
        uint64_t S2 = 0, T3 = 0;
        uint32_t k = 0;
        if (wide) {
            const uint32_t* row = pc.w + x * rs;
            for (uint32_t d = 0; d < NB; ++d) {
                const uint32_t cnt = row[d];
                const uint2 be = bins[d];
                S2 += (uint64_t)cnt * be.x;
                const uint32_t t = be.y + cnt;
                const uint64_t t2 = (uint64_t)t * t;
                T3 += t2 * t;
                k += cnt;
            }
        } else {
            auto bin = [&](uint32_t d, uint32_t cnt) {
                const uint2 be = bins[d];
                S2 += (uint64_t)cnt * be.x;
                const uint32_t t = be.y + cnt;
                T3 += (uint64_t)(uint32_t)__umul24(t, t) * t;
                k += cnt;
            };
            // rows start on an even counter: two counters per word
            const uint32_t* row = pc.w + ((x * rs) >> 1);
            uint32_t d = 0;
            for (; d + 4u <= NB; d += 4u) {
                const uint32_t w0 = row[d >> 1], w1 = row[(d >> 1) + 1u];
                bin(d, w0 & 0xFFFFu);
                bin(d + 1u, w0 >> 16);
                bin(d + 2u, w1 & 0xFFFFu);
                bin(d + 3u, w1 >> 16);
            }
            for (; d < NB; d += 2u) {
                const uint32_t cw = row[d >> 1];
                bin(d, cw & 0xFFFFu);
                if (d + 1u < NB) bin(d + 1u, cw >> 16);
            }
        }
        acc[2u * x] = S2;
        acc[2u * x + 1u] = T3 - (uint64_t)(na + k) - ta;
        cnz[x] = k;
    }
    if (tid == 0) cnz[0] = na;
    __syncthreads();

    // reference values below the zero block
    uint32_t ltz0;
    if (mode == kZeroMin) ltz0 = 0u;
    else if (mode == kZeroMax) ltz0 = na;
    else if (packed) ltz0 = vmin == 0 ? 0u : ((uint32_t)vmin < NB ? st.scan[16] : na);
    else ltz0 = vmin >= 0 ? 0u : ((uint32_t)(-vmin) < NB ? st.scan[16] : na);

    // per row: errors of this column (mannwhitneyu.cpp:936-1038); per target
    // row: zero block, U1, tie, sizes, asymptotic p
    const uint64_t ebase = (uint64_t)c * (uint64_t)(2u * G + 1u);
    for (uint32_t x = tid; x < rows; x += nt) {
        const uint32_t g = x == 0u ? 0u : t0 + x;
        const uint32_t nz = cnz[x], ez = cez[x], gr = grs[x];
        if ((uint64_t)nz + ez > gr) atomicMin(J.err_key, (unsigned long long)(ebase + g));
        const uint32_t ns = (mode == kZeroNone) ? nz + ez : gr;
        if (ns < J.min_n) atomicMin(J.err_key, (unsigned long long)(ebase + (g == 0u ? G : G + g)));
        if (x == 0u) continue;
        uint32_t zb, n1, n2;
        if (mode == kZeroNone) {
            zb = ez;
            n1 = na + za;
            n2 = nz + zb;
        } else {
            zb = gr > nz ? gr - nz : 0u;
            n1 = gr0;
            n2 = gr;
        }
        const uint64_t S2 = acc[2u * x] + (uint64_t)zb * (2ull * ltz0 + za);
        const uint64_t tie = acc[2u * x + 1u] + ta + tie_f((uint64_t)za + zb);
        const uint64_t twoN = 2ull * (uint64_t)n1 * (uint64_t)n2;
        const uint64_t u1x2 = twoN >= S2 ? twoN - S2 : 0ull;
        const int64_t idx = c * J.T + (int64_t)(g - 1u);
        const double U1 = (double)u1x2 * 0.5;
        const double td = J.tie_corr ? (double)tie : 0.0;
        gl(J.U1)[idx] = U1;
        if (J.tie) gl(J.tie)[idx] = td;
        if (J.n1o) { gl(J.n1o)[idx] = (double)n1; gl(J.n2o)[idx] = (double)n2; }
        if (J.asym) gl(J.P)[idx] = p_asymptotic(U1, (double)n1, (double)n2, td, J.cc, J.alt);
    }
}

// The plain launch: workgroup b takes item b of one job.
template <typename V, int NT>
__global__ __launch_bounds__(NT) void count_kernel(const JobDesc J,
                                                   const CountItem* __restrict__ items,
                                                   uint32_t lds_bytes) {
    __shared__ CountState st;
    count_item<V, NT>(J, items[blockIdx.x], lds_bytes, st);
}

// The batched launch: the items of several jobs (each names its job) in one
// queue ordered largest first, one workgroup per item.  The hardware hands
// out workgroups in queue order as CUs free up, so long columns start first
// and short ones of any job fill in behind them -- the queue is drained by
// the dispatcher rather than by a software loop (an item loop inside the
// kernel keeps values of one item live across items: 178 to 256 VGPRs
// against 106, one or two waves per SIMD instead of four).
template <typename V, int NT>
__global__ __launch_bounds__(NT) void count_queue_kernel(const JobDesc* __restrict__ jobs,
                                                         const CountItem* __restrict__ items,
                                                         uint32_t lds_bytes) {
    __shared__ CountState st;
    const CountItem it = items[blockIdx.x];
    const uint32_t job = uniform_u32(it.flags >> kCountJobShift);
    count_item<V, NT>(jobs[job], it, lds_bytes, st);
}

// Metadata of a resident matrix, computed once at upload: ok[c] = 1 when every
// stored value of column c is non-finite or an integer in [-32768, 32767];
// range[2c], range[2c+1] = smallest and largest stored nonzero finite value
// (range[2c] > range[2c+1] when there is none).
template <typename V>
__global__ __launch_bounds__(256) void classify_columns_kernel(const V* vals,
                                                               const int64_t* indptr,
                                                               uint8_t* ok_out, int32_t* range,
                                                               const int32_t* rows, uint32_t R,
                                                               uint8_t* ident_out) {
    __shared__ uint32_t ok, ident;
    __shared__ int32_t lo, hi;
    const int64_t c = blockIdx.x;
    if (threadIdx.x == 0) { ok = 1u; lo = INT_MAX; hi = INT_MIN; ident = 1u; }
    __syncthreads();
    // a column storing exactly rows 0 .. R-1 in order (a dense matrix held as
    // CSC): its readers can take the row from the position
    const int64_t q0 = indptr[c], q1 = indptr[c + 1];
    if (q1 - q0 == (int64_t)R) {
        bool in_order = true;
        for (int64_t p = q0 + threadIdx.x; p < q1; p += blockDim.x)
            in_order &= rows[p] == (int32_t)(p - q0);
        if (!in_order) atomicAnd(&ident, 0u);
    } else if (threadIdx.x == 0) {
        ident = 0u;
    }
    bool li = true;
    int mn = INT_MAX, mx = INT_MIN;
    for (int64_t p = indptr[c] + threadIdx.x; p < indptr[c + 1]; p += blockDim.x) {
        const V v = vals[p];
        if (!KeyOf<V>::finite(v) || v == V(0)) continue;
        int iv;
        if (KeyOf<V>::small_int(v, iv) && iv >= -32768 && iv <= 32767) {
            mn = min(mn, iv);
            mx = max(mx, iv);
        } else {
            li = false;
        }
    }
    if (!li) atomicAnd(&ok, 0u);
    atomicMin(&lo, mn);
    atomicMax(&hi, mx);
    __syncthreads();
    if (threadIdx.x == 0) {
        ok_out[c] = (uint8_t)ok;
        range[2 * c] = ok ? lo : 1;
        range[2 * c + 1] = ok ? hi : 0;
        ident_out[c] = (uint8_t)ident;
    }
}

// Packed entries (pack_row_bits) of the columns with lo[c] != INT_MIN, one
// workgroup per column.  Classification found every finite nonzero value of
// such a column an integer in [lo, hi] = range[2c .. 2c+1], and the host
// checked that hi - lo + 1 leaves the two special codes free.  The code of a
// value is its rank among the column's distinct nonzero values, so the count
// kernel's bins are only the values that occur: a first pass marks them in an
// LDS bitmap (at most 2^16 values: 2,048 words), a prefix of the word
// popcounts gives the ranks, a second pass writes the words.  dense[c] = the
// distinct values, zneg[c] = those below zero.
template <typename V>
__global__ __launch_bounds__(256) void pack_columns_kernel(const V* vals, const int32_t* rows,
                                                           const int64_t* indptr,
                                                           const int32_t* lo,
                                                           const int32_t* range, uint32_t R,
                                                           uint32_t vb, uint32_t* out,
                                                           uint32_t* dense, uint32_t* zneg) {
    __shared__ uint32_t bits[kPackBitmapWords];
    __shared__ uint32_t pre[kPackBitmapWords];
    __shared__ uint32_t scan[20];
    const int64_t c = blockIdx.x;
    const int32_t base = lo[c];
    if (base == INT_MIN) return;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int64_t span = (int64_t)range[2 * c + 1] - base + 1;   // 0 or less: no nonzero value
    const uint32_t words = span > 0 ? (uint32_t)((span + 31) / 32) : 0u;
    for (uint32_t w = tid; w < words; w += nt) bits[w] = 0u;
    __syncthreads();
    const int64_t p0 = indptr[c], p1 = indptr[c + 1];
    for (int64_t p = p0 + tid; p < p1; p += nt) {
        const V v = vals[p];
        if (KeyOf<V>::finite(v) && v != V(0)) {
            const uint32_t k = (uint32_t)((int32_t)v - base);
            atomicOr(&bits[k >> 5], 1u << (k & 31u));
        }
    }
    __syncthreads();
    for (uint32_t w = tid; w < words; w += nt) pre[w] = (uint32_t)__builtin_popcount(bits[w]);
    __syncthreads();
    const uint32_t total = block_excl_scan(pre, (int)words, scan);
    // rank of offset k: present values at smaller offsets
    auto rank = [&](uint32_t k) {
        return pre[k >> 5] + (uint32_t)__builtin_popcount(bits[k >> 5] & ((1u << (k & 31u)) - 1u));
    };
    if (tid == 0) {
        dense[c] = total;
        zneg[c] = base >= 0 ? 0u : ((int64_t)(-(int64_t)base) < span ? rank((uint32_t)(-base)) : total);
    }
    // an explicit zero takes the code one past the last bin (the count
    // kernel's zero column: a column with no nonzero value has one bin)
    const uint32_t zc = total > 0u ? total : 1u, sc = pack_skip_code(vb);
    for (int64_t p = p0 + tid; p < p1; p += nt) {
        const V v = vals[p];
        const int32_t r = rows[p];
        const uint32_t row = (r >= 0 && (uint32_t)r < R) ? (uint32_t)r : R;
        uint32_t code;
        if (!KeyOf<V>::finite(v)) code = sc;
        else if (v == V(0)) code = zc;
        else code = rank((uint32_t)((int32_t)v - base));
        out[p] = (row << vb) | code;
    }
}

__global__ void init_job_state_kernel(unsigned long long* err_key, uint32_t* status,
                                      unsigned long long* status_col) {
    if (threadIdx.x == 0) {
        *err_key = ~0ull;
        *status = 0u;
        *status_col = ~0ull;
    }
}

// ---------------------------------------------------------------------------
// Exact null tables (one workgroup per distinct (n1, n2)), FP64, the same
// recurrence and the same order of additions as the oracle restatement of
// src/hpdex/backend/pykernel_.py:173-185, so every count is bit-identical.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void exact_table_kernel(const ExactTable* tabs, double* tabbuf,
                                                           double* scratch,
                                                           size_t scratch_per_table) {
    const ExactTable t = tabs[blockIdx.x];
    const int64_t n1 = t.n1, n2 = t.n2;
    const int64_t s = n1 < n2 ? n1 : n2;
    const int64_t l = n1 < n2 ? n2 : n1;
    const int64_t U = n1 * n2, W = U + 1;
    double* fp = scratch + (size_t)blockIdx.x * scratch_per_table;
    double* fc = fp + (size_t)(s + 1) * W;
    double* out = tabbuf + t.offset;
    const int tid = tid_x(), nt = blockDim.x;
    for (int64_t x = tid; x < (s + 1) * W; x += nt) fp[x] = (x % W == 0) ? 1.0 : 0.0;
    __syncthreads();
    for (int64_t i = 1; i <= l; ++i) {
        const int64_t cap = i * s;
        for (int64_t k = tid; k < W; k += nt) {
            double below = (k == 0) ? 1.0 : 0.0;
            fc[k] = below;
            for (int64_t j = 1; j <= s; ++j) {
                double val = 0.0;
                if (k <= cap) val = (k >= j) ? below + fp[j * W + k - j] : below;
                fc[j * W + k] = val;
                below = val;
            }
        }
        __syncthreads();
        double* tmp = fp; fp = fc; fc = tmp;
    }
    const double* cnt = fp + (size_t)s * W;
    for (int64_t k = tid; k < W; k += nt) out[1 + k] = cnt[k];
    // sequential prefix (same order as a left-to-right loop) by the first wave
    if (tid < kWave) {
        double acc = 0.0;
        for (int64_t b = 0; b < W; b += kWave) {
            const int64_t k = b + tid;
            const double x = k < W ? cnt[k] : 0.0;
            double mine = 0.0;
            const int lim = (int)((W - b) < kWave ? (W - b) : kWave);
            for (int l2 = 0; l2 < lim; ++l2) {
                acc += __shfl(x, l2, kWave);
                if (tid == l2) mine = acc;
            }
            if (k < W) out[1 + W + k] = mine;
        }
        if (tid == 0) out[0] = acc;
    }
}

__global__ void exact_p_kernel(ExactJob E) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= E.n_pairs) return;
    const int32_t ti = E.tidx_by_g ? E.tidx_by_g[i % E.T] : E.tidx_by_pair[i];
    const ExactTable tb = E.tables[ti];
    const int64_t n1 = tb.n1, n2 = tb.n2, Umax = n1 * n2, W = Umax + 1;
    const double* base = E.tabbuf + tb.offset;
    const double U1 = E.U1[i];
    const double nm = (double)n1 * (double)n2;
    double Ustat;
    if (E.alt == 2) { const double U2 = nm - U1; Ustat = U1 > U2 ? U1 : U2; }
    else if (E.alt == 1) Ustat = U1;
    else Ustat = nm - U1;
    double sf = 1.0;
    if (Umax > 0) {
        const double Uc = Ustat < 0.0 ? 0.0 : (Ustat > (double)Umax ? (double)Umax : Ustat);
        const int64_t u = (int64_t)floor(Uc);
        const int64_t kc = Umax - u;
        const int64_t small = u < kc ? u : kc;
        const double total = base[0];
        const double cdf_small = base[1 + W + small];
        const double pmf_small = base[1 + small];
        sf = (u <= kc) ? (1.0 - cdf_small / total + pmf_small / total) : (cdf_small / total);
    }
    double p = (E.alt == 2) ? 2.0 * sf : sf;
    if (p < 0.0) p = 0.0;
    if (p > 1.0) p = 1.0;
    E.P[i] = p;
}

// ---------------------------------------------------------------------------
// group_mean (src/cpp/mannwhitneyu.cpp:1118-1192): one workgroup per column,
// per-group FP64 sums and counts in LDS.
// ---------------------------------------------------------------------------
// Workgroup (c, y) takes column c and groups [y*S, y*S + S): with more groups
// than one workgroup's LDS holds accumulators for, each slice of S groups
// reads the column once.
template <typename V>
__global__ __launch_bounds__(1024) void group_mean_kernel(const V* vals, const int32_t* rows,
                                                          const int64_t* indptr,
                                                          const int32_t* gid,
                                                          const uint32_t* group_rows,
                                                          int32_t R, int32_t G, int32_t S,
                                                          int include_zeros, double* out) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int32_t g0 = (int32_t)blockIdx.y * S;
    const uint32_t gs = (uint32_t)min(S, G - g0);
    double* sum = (double*)smem;
    uint32_t* vc = (uint32_t*)(sum + gs);
    uint32_t* ic = vc + gs;
    const int64_t c = blockIdx.x;
    const int tid = tid_x(), nt = blockDim.x;
    for (uint32_t j = tid; j < gs; j += nt) { sum[j] = 0.0; vc[j] = 0u; ic[j] = 0u; }
    __syncthreads();
    const int64_t p0 = indptr[c], p1 = indptr[c + 1];
    for_each_entry<V>(rows, vals, gid, (uint32_t)R, (uint32_t)G, p0, p1, [&](uint32_t g, V x) {
        const uint32_t j = g - (uint32_t)g0;
        if (j >= gs) return;
        const double v = (double)x;
        if (__builtin_isfinite(v)) { atomicAdd(&sum[j], v); atomicAdd(&vc[j], 1u); }
        else atomicAdd(&ic[j], 1u);
    });
    __syncthreads();
    for (uint32_t j = tid; j < gs; j += nt) {
        const uint32_t g = (uint32_t)g0 + j;
        const uint64_t denom = include_zeros ? ((uint64_t)group_rows[g] - (uint64_t)ic[j])
                                             : (uint64_t)vc[j];
        out[c * G + g] = denom > 0 ? sum[j] / (double)denom : 0.0;
    }
}

// ---------------------------------------------------------------------------
// Upload conversions.
// ---------------------------------------------------------------------------
__global__ void rows_i64_to_i32_kernel(const int64_t* in, int32_t* out, size_t n, int64_t R) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int64_t r = in[i];
        out[i] = (r >= 0 && r < R) ? (int32_t)r : -1;
    }
}

template <typename Tin, typename Tout>
__global__ void convert_values_kernel(const Tin* in, Tout* out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        out[i] = (Tout)in[i];
}

// Streaming copy used to calibrate HBM counters and measure attainable
// bandwidth: 16-byte loads and stores, grid-stride.
__global__ __launch_bounds__(256) void stream_copy_kernel(const uint4* __restrict__ in,
                                                          uint4* __restrict__ out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        out[i] = in[i];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned grid_for(size_t n, int block) {
    size_t g = (n + block - 1) / block;
    if (g > 65536) g = 65536;
    return (unsigned)(g ? g : 1);
}

hipError_t launch_rows_i64_to_i32(const int64_t* in, int32_t* out, size_t n, int64_t R,
                                  hipStream_t s) {
    if (n == 0) return hipSuccess;
    rows_i64_to_i32_kernel<<<grid_for(n, 256), 256, 0, s>>>(in, out, n, R);
    return hipGetLastError();
}

bool dtype_uses_double(int dtype) {
    // int8/uint8/int16/uint16/float32 convert to float exactly; the rest to double
    return !(dtype == 0 || dtype == 1 || dtype == 2 || dtype == 3 || dtype == 8);
}

hipError_t launch_convert_values(int dtype, const void* in, void* out, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const unsigned gr = grid_for(n, 256);
    switch (dtype) {
    case 0: convert_values_kernel<int8_t, float><<<gr, 256, 0, s>>>((const int8_t*)in, (float*)out, n); break;
    case 1: convert_values_kernel<uint8_t, float><<<gr, 256, 0, s>>>((const uint8_t*)in, (float*)out, n); break;
    case 2: convert_values_kernel<int16_t, float><<<gr, 256, 0, s>>>((const int16_t*)in, (float*)out, n); break;
    case 3: convert_values_kernel<uint16_t, float><<<gr, 256, 0, s>>>((const uint16_t*)in, (float*)out, n); break;
    case 4: convert_values_kernel<int32_t, double><<<gr, 256, 0, s>>>((const int32_t*)in, (double*)out, n); break;
    case 5: convert_values_kernel<uint32_t, double><<<gr, 256, 0, s>>>((const uint32_t*)in, (double*)out, n); break;
    case 6: convert_values_kernel<int64_t, double><<<gr, 256, 0, s>>>((const int64_t*)in, (double*)out, n); break;
    case 7: convert_values_kernel<uint64_t, double><<<gr, 256, 0, s>>>((const uint64_t*)in, (double*)out, n); break;
    default: return hipErrorInvalidValue;   // float32/float64 are copied as-is
    }
    return hipGetLastError();
}

static inline unsigned align16(size_t b) { return (unsigned)((b + 15) & ~(size_t)15); }

// LDS of the reference-table and target kernels: the larger of a full
// histogram table and the on-chip sorted reference plus its bucket index.
static unsigned table_lds_bytes(bool f64) { return ref_table_lds(f64); }

// The target phase of the sort-path columns: 4-byte keys take the window
// pass, then the sorted pass over the columns it flagged; 8-byte keys the
// sorted pass over every column.
template <typename V>
static void launch_target_pass(const JobDesc* jobs, const WorkItem* items, uint32_t n,
                               hipStream_t s) {
    if constexpr (sizeof(V) == 4)
        target_window_kernel<V><<<n, kTargetBlock, target_lds(false), s>>>(jobs, items, n);
    target_kernel<V><<<n, kColBlock, table_lds_bytes(sizeof(V) == 8), s>>>(jobs, items, n);
}

hipError_t prepare_count_kernel(uint32_t bytes) {
    const void* fns[] = {(const void*)count_kernel<float, 256>, (const void*)count_kernel<float, 512>,
                         (const void*)count_kernel<float, 1024>,
                         (const void*)count_kernel<double, 256>,
                         (const void*)count_kernel<double, 512>,
                         (const void*)count_kernel<double, 1024>,
                         (const void*)count_queue_kernel<float, 256>,
                         (const void*)count_queue_kernel<float, 512>,
                         (const void*)count_queue_kernel<float, 1024>,
                         (const void*)count_queue_kernel<double, 256>,
                         (const void*)count_queue_kernel<double, 512>,
                         (const void*)count_queue_kernel<double, 1024>};
    for (const void* f : fns) {
        const hipError_t e =
            hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (e != hipSuccess) return e;
    }
    const void* split_fns[] = {(const void*)split_kernel<float>, (const void*)split_kernel<double>};
    for (const void* f : split_fns) {
        const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 (int)kSplitLdsMax);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_mwu_pipeline(const PipelineLaunch& L, hipStream_t s, hipEvent_t* ev) {
    if (L.n_count <= 0 && L.n_gen <= 0) return hipSuccess;
    auto mark = [&](int i) { if (ev) (void)hipEventRecord(ev[i], s); };
    mark(0);
    if (L.n_count > 0) {
#define HPDEX_COUNT(V, NT) \
    count_kernel<V, NT><<<L.n_count, NT, L.count_lds, s>>>(L.job, L.count_items, L.count_lds)
        if (L.f64) {
            if (L.count_block == 256) HPDEX_COUNT(double, 256);
            else if (L.count_block == 1024) HPDEX_COUNT(double, 1024);
            else HPDEX_COUNT(double, 512);
        } else {
            if (L.count_block == 256) HPDEX_COUNT(float, 256);
            else if (L.count_block == 1024) HPDEX_COUNT(float, 1024);
            else HPDEX_COUNT(float, 512);
        }
#undef HPDEX_COUNT
    }
    mark(1);
    if (L.n_gen > 0) {
        // the most LDS the split kernel may take: one 1024-thread workgroup
        // per CU fits in any case, and a column whose key range fits in what
        // is left gathers its keys there and writes them out coalesced; with
        // more groups than that LDS holds counters for, the split counts in HBM
        const unsigned split_lds = (unsigned)L.split_lds_max;
        const bool global_split = split_counter_bytes((uint64_t)L.G) > (uint64_t)split_lds;
        const unsigned tab_lds = table_lds_bytes(L.f64);
        const unsigned grid = (unsigned)L.n_gen, n = (uint32_t)L.n_gen;
        auto sort_path = [&](auto v_tag) {
            using V = typename decltype(v_tag)::type;
            if (global_split)
                split_global_kernel<V><<<grid, kColBlock, 0, s>>>(L.jobs, L.gen_items, n);
            else
                split_kernel<V><<<grid, kSplitBlock, split_lds, s>>>(L.jobs, L.gen_items, n,
                                                                     split_lds);
            mark(2);
            ref_kernel<V><<<grid, kRefBlock, tab_lds, s>>>(L.jobs, L.gen_items, n);
            mark(3);
            launch_target_pass<V>(L.jobs, L.gen_items, n, s);
        };
        if (L.f64) sort_path(TypeTag<double>{});
        else sort_path(TypeTag<float>{});
    } else {
        mark(2);
        mark(3);
    }
    mark(4);
    return hipGetLastError();
}

// One stage of the launch sequence alone (0 count, 1 split, 2 reference, 3
// target: the window and sorted passes), for timing back-to-back launches of
// it.  Every stage may run again on the outputs of the sequence before it:
// each writes what it wrote the first time.
hipError_t launch_mwu_stage(const PipelineLaunch& L, int stage, hipStream_t s) {
    if (stage == 0) {
        if (L.n_count <= 0) return hipSuccess;
#define HPDEX_COUNT(V, NT) \
    count_kernel<V, NT><<<L.n_count, NT, L.count_lds, s>>>(L.job, L.count_items, L.count_lds)
        if (L.f64) {
            if (L.count_block == 256) HPDEX_COUNT(double, 256);
            else if (L.count_block == 1024) HPDEX_COUNT(double, 1024);
            else HPDEX_COUNT(double, 512);
        } else {
            if (L.count_block == 256) HPDEX_COUNT(float, 256);
            else if (L.count_block == 1024) HPDEX_COUNT(float, 1024);
            else HPDEX_COUNT(float, 512);
        }
#undef HPDEX_COUNT
        return hipGetLastError();
    }
    if (L.n_gen <= 0) return hipSuccess;
    const unsigned split_lds = (unsigned)L.split_lds_max;
    const bool global_split = split_counter_bytes((uint64_t)L.G) > (uint64_t)split_lds;
    const unsigned tab_lds = table_lds_bytes(L.f64);
    const unsigned grid = (unsigned)L.n_gen, n = (uint32_t)L.n_gen;
    auto one = [&](auto v_tag) {
        using V = typename decltype(v_tag)::type;
        if (stage == 1) {
            if (global_split)
                split_global_kernel<V><<<grid, kColBlock, 0, s>>>(L.jobs, L.gen_items, n);
            else
                split_kernel<V><<<grid, kSplitBlock, split_lds, s>>>(L.jobs, L.gen_items, n,
                                                                     split_lds);
        } else if (stage == 2) {
            ref_kernel<V><<<grid, kRefBlock, tab_lds, s>>>(L.jobs, L.gen_items, n);
        } else {
            launch_target_pass<V>(L.jobs, L.gen_items, n, s);
        }
    };
    if (L.f64) one(TypeTag<double>{});
    else one(TypeTag<float>{});
    return hipGetLastError();
}

hipError_t launch_mwu_batch(const BatchLaunch& L, hipStream_t s) {
    if (L.n_count > 0) {
#define HPDEX_COUNTQ(V, NT)                                                                  \
    count_queue_kernel<V, NT><<<(unsigned)L.n_count, NT, L.count_lds, s>>>(L.jobs, L.count_items, \
                                                                          L.count_lds)
        if (L.f64) {
            if (L.count_block == 256) HPDEX_COUNTQ(double, 256);
            else if (L.count_block == 1024) HPDEX_COUNTQ(double, 1024);
            else HPDEX_COUNTQ(double, 512);
        } else {
            if (L.count_block == 256) HPDEX_COUNTQ(float, 256);
            else if (L.count_block == 1024) HPDEX_COUNTQ(float, 1024);
            else HPDEX_COUNTQ(float, 512);
        }
#undef HPDEX_COUNTQ
    }
    if (L.n_gen > 0) {
        const unsigned split_lds = (unsigned)L.split_lds_max;
        const bool global_split = split_counter_bytes((uint64_t)L.G) > (uint64_t)split_lds;
        const unsigned tab_lds = table_lds_bytes(L.f64);
        const unsigned grid = (unsigned)L.n_gen, n = (uint32_t)L.n_gen;
        auto sort_path = [&](auto v_tag) {
            using V = typename decltype(v_tag)::type;
            if (global_split)
                split_global_kernel<V><<<grid, kColBlock, 0, s>>>(L.jobs, L.gen_items, n);
            else
                split_kernel<V><<<grid, kSplitBlock, split_lds, s>>>(L.jobs, L.gen_items, n,
                                                                     split_lds);
            ref_kernel<V><<<grid, kRefBlock, tab_lds, s>>>(L.jobs, L.gen_items, n);
            launch_target_pass<V>(L.jobs, L.gen_items, n, s);
        };
        if (L.f64) sort_path(TypeTag<double>{});
        else sort_path(TypeTag<float>{});
    }
    return hipGetLastError();
}

hipError_t launch_refcache_fill(const JobDesc* job, bool f64, const int32_t* cols, uint32_t n,
                                RefCacheCol* rc, void* rc_keys, uint32_t* rc_tab, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (f64)
        refcache_fill_kernel<uint64_t><<<n, kColBlock, 0, s>>>(job, cols, n, rc,
                                                                (uint64_t*)rc_keys, rc_tab);
    else
        refcache_fill_kernel<uint32_t><<<n, kColBlock, 0, s>>>(job, cols, n, rc,
                                                                (uint32_t*)rc_keys, rc_tab);
    return hipGetLastError();
}

hipError_t launch_classify_columns(const void* vals, bool f64, const int64_t* indptr, int64_t C,
                                   uint8_t* ok, int32_t* range, const int32_t* rows, uint32_t R,
                                   uint8_t* ident, hipStream_t s) {
    if (C <= 0) return hipSuccess;
    if (f64)
        classify_columns_kernel<double><<<(unsigned)C, 256, 0, s>>>(
            (const double*)vals, indptr, ok, range, rows, R, ident);
    else
        classify_columns_kernel<float><<<(unsigned)C, 256, 0, s>>>(
            (const float*)vals, indptr, ok, range, rows, R, ident);
    return hipGetLastError();
}

hipError_t launch_pack_columns(const void* vals, bool f64, const int32_t* rows,
                               const int64_t* indptr, int64_t C, const int32_t* lo,
                               const int32_t* range, uint32_t R, uint32_t vb, uint32_t* out,
                               uint32_t* dense, uint32_t* zneg, hipStream_t s) {
    if (C <= 0) return hipSuccess;
    if (f64)
        pack_columns_kernel<double><<<(unsigned)C, 256, 0, s>>>(
            (const double*)vals, rows, indptr, lo, range, R, vb, out, dense, zneg);
    else
        pack_columns_kernel<float><<<(unsigned)C, 256, 0, s>>>(
            (const float*)vals, rows, indptr, lo, range, R, vb, out, dense, zneg);
    return hipGetLastError();
}

hipError_t launch_stream_copy(const void* in, void* out, size_t bytes, int grid, hipStream_t s) {
    const size_t n = bytes / 16;
    if (n == 0) return hipSuccess;
    stream_copy_kernel<<<grid, 256, 0, s>>>((const uint4*)in, (uint4*)out, n);
    return hipGetLastError();
}

hipError_t launch_init_job_state(unsigned long long* err_key, uint32_t* status,
                                 unsigned long long* status_col, hipStream_t s) {
    init_job_state_kernel<<<1, 64, 0, s>>>(err_key, status, status_col);
    return hipGetLastError();
}

hipError_t launch_exact_tables(const ExactTable* tabs, int n_tabs, double* tabbuf,
                               double* scratch, size_t scratch_per_table, hipStream_t s) {
    if (n_tabs <= 0) return hipSuccess;
    exact_table_kernel<<<n_tabs, 1024, 0, s>>>(tabs, tabbuf, scratch, scratch_per_table);
    return hipGetLastError();
}

hipError_t launch_exact_p(const ExactJob& E, hipStream_t s) {
    if (E.n_pairs <= 0) return hipSuccess;
    const unsigned gr = (unsigned)((E.n_pairs + 255) / 256);
    exact_p_kernel<<<gr, 256, 0, s>>>(E);
    return hipGetLastError();
}

hipError_t launch_group_mean(const void* vals, bool f64, const int32_t* rows,
                             const int64_t* indptr, const int32_t* gid,
                             const uint32_t* group_rows, int64_t C, int32_t R, int32_t G,
                             int include_zeros, double* out, hipStream_t s) {
    if (C <= 0 || G <= 0) return hipSuccess;
    // 16 B of accumulators per group: slices of at most 4,096 groups
    const int32_t S = G < (int32_t)(kLdsBudget / 16) ? G : (int32_t)(kLdsBudget / 16);
    const dim3 grid((unsigned)C, (unsigned)((G + S - 1) / S));
    const size_t need = 16u * (size_t)S;
    if (f64)
        group_mean_kernel<double><<<grid, kColBlock, align16(need), s>>>(
            (const double*)vals, rows, indptr, gid, group_rows, R, G, S, include_zeros, out);
    else
        group_mean_kernel<float><<<grid, kColBlock, align16(need), s>>>(
            (const float*)vals, rows, indptr, gid, group_rows, R, G, S, include_zeros, out);
    return hipGetLastError();
}

}  // namespace hpdex
