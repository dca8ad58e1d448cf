// hpdex_kernels.hip -- gfx950 kernels of the Mann-Whitney U backend.
//
// The per-gene kernel fuses what the reference does per column
// (src/cpp/mannwhitneyu.cpp:903-1043): group split, sort of the reference
// slice once per gene, U and tie term of every target against it, and the
// asymptotic p-value (:191-240).  It never materialises a merge: for each
// target value v it needs only lt(v) = #reference values below v and
// eq(v) = #reference values equal to v, so
//
//   2*U_target = sum_b (2*lt(b) + eq(b))            over all target values b
//   U1         = n1*n2 - U_target                    (reference-group U)
//   tie        = sum over distinct values of t^3 - t, built per element as
//                3*x*(x-1) with x = eq(b) + j(b), j(b) = 1-based position of b
//                inside its run of equal target values, plus the reference's
//                own runs and the zero block.
//
// since f(a+b) - f(a) = sum_{j=1..b} 3(a+j)(a+j-1) for f(t) = t^3 - t.  All of
// it is exact integer arithmetic (U held as 2U in 64 bits, ties in 64 bits),
// which is what makes U1 and the tie term bit-identical to the reference's
// midrank merge (:618-840).
//
// Zero handling (:935-1039): zeros are never stored.  Explicit zeros, and in
// min/max/mix also absent and non-finite entries, form one block whose
// position is below every value (min), above every value (max) or at 0
// (mix, and none -- where only explicit zeros count).
#include <climits>

#include "hpdex_device.h"
#include "hpdex_internal.h"

namespace hpdex {

constexpr int kChunk = 1024;          // largest target run sorted in registers
constexpr int kMaxBuckets = 16384;    // bucket index over the reference keys

struct ColState {
    uint64_t kmin, kmax;   // reference key range (sort path)
    uint64_t TA;           // reference-only tie sum over nonzero values
    uint32_t all_int;      // every in-group nonzero value is a small integer
    int32_t vmin, vmax;    // their range
    uint32_t refneg;       // reference nonzero values below zero
    uint32_t refcur;
    uint32_t next_t;
    uint32_t hist;         // 1 = histogram path, 0 = sort path
    uint32_t nb;           // histogram bins or reference buckets
    uint32_t shift;        // bucket = (key - kmin) >> shift
    uint32_t P;            // padded reference length (sort path)
    uint32_t tgt_cap;      // target keys that fit the LDS region
    uint32_t glo, ghi, off0;
    uint32_t fail;
    uint32_t ttotal;       // reused as the "a target is too large" flag
    int32_t item;
    uint32_t scan[20];
    uint64_t red[16];
};

__device__ inline uint64_t tie_f(uint64_t t) { return t * t * t - t; }

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
template <typename K>
__device__ inline uint32_t lower_bound_lds(const K* a, uint32_t lo, uint32_t hi, K v) {
    uint32_t n = hi - lo;
    while (n > 0) {
        const uint32_t h = n >> 1;
        if (a[lo + h] < v) { lo += h + 1; n -= h + 1; } else { n = h; }
    }
    return lo;
}
// first index in a[lo, hi) with a[i] > v
template <typename K>
__device__ inline uint32_t upper_bound_lds(const K* a, uint32_t lo, uint32_t hi, K v) {
    uint32_t n = hi - lo;
    while (n > 0) {
        const uint32_t h = n >> 1;
        if (!(v < a[lo + h])) { lo += h + 1; n -= h + 1; } else { n = h; }
    }
    return lo;
}

// Reference table seen by the target phase.
template <typename K>
struct RefView {
    const uint32_t* cum;   // histogram path: exclusive prefix of bin counts
    const K* keys;         // sort path: sorted reference keys
    const uint32_t* bidx;  // sort path: bucket -> first key index
    uint64_t kmin, kmax;
    uint32_t shift, na, hist;

    __device__ inline void lookup(K v, uint32_t& lt, uint32_t& eq) const {
        if (hist) {
            const uint32_t a = cum[(uint32_t)v], b = cum[(uint32_t)v + 1];
            lt = a; eq = b - a;
            return;
        }
        if (na == 0 || (uint64_t)v < kmin) { lt = 0; eq = 0; return; }
        if ((uint64_t)v > kmax) { lt = na; eq = 0; return; }
        const uint32_t bkt = (uint32_t)(((uint64_t)v - kmin) >> shift);
        const uint32_t lo = bidx[bkt], hi = bidx[bkt + 1];
        const uint32_t l = lower_bound_lds<K>(keys, lo, hi, v);
        const uint32_t u = upper_bound_lds<K>(keys, l, hi, v);
        lt = l; eq = u - l;
    }
};

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
__device__ inline bool zero_below(const TargetCtx<typename KeyOf<V>::K>& t, typename KeyOf<V>::K v) {
    if (t.mode == kZeroMin) return true;
    if (t.mode == kZeroMax) return false;
    if (t.ref.hist) return (int32_t)v + t.vmin > 0;
    return v >= KeyOf<V>::kPosBit;
}

// Sort one run of up to 64*KK target keys in registers and accumulate its
// contribution.  seg: the target's keys in LDS; q: chunk number; earlier
// chunks (already sorted in place) are searched for equal keys so that j
// counts position inside the whole run of equal values.
template <typename V, int KK>
__device__ __forceinline__ void target_chunk(TargetCtx<typename KeyOf<V>::K>& t, typename KeyOf<V>::K* seg,
                             uint32_t m, int q, bool store) {
    using K = typename KeyOf<V>::K;
    const int lane = lane_id();
    K* base = seg + (size_t)q * kChunk;
    K v[KK];
#pragma unroll
    for (int i = 0; i < KK; ++i) {
        const uint32_t e = (uint32_t)(i * kWave + lane);
        v[i] = e < m ? base[e] : KeyOf<V>::kSentinel;
    }
    wave_sort<KK, K>(v);
    if (store) {
#pragma unroll
        for (int i = 0; i < KK; ++i) {
            const uint32_t e = (uint32_t)(i * kWave + lane);
            if (e < m) base[e] = v[i];
        }
        __threadfence_block();
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
                const uint32_t l = lower_bound_lds<K>(pc, 0, kChunk, v[i]);
                const uint32_t u = upper_bound_lds<K>(pc, l, kChunk, v[i]);
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

template <typename V>
__device__ __forceinline__ void process_target(const JobDesc& J, int64_t c, uint32_t g, ColState& st,
                               const uint32_t* cnt_nz, const uint32_t* cnt_ez, const uint32_t* seg,
                               typename KeyOf<V>::K* tgt, const RefView<typename KeyOf<V>::K>& ref) {
    using K = typename KeyOf<V>::K;
    const int lane = lane_id();
    const uint32_t nzb = cnt_nz[g];
    const uint32_t start = seg[g] - nzb - st.off0;
    const uint32_t nza = cnt_nz[0];
    const uint32_t gr0 = J.group_rows[0], grg = J.group_rows[g];
    uint32_t za, zb, n1, n2;
    if (J.mode == kZeroNone) {
        za = cnt_ez[0]; zb = cnt_ez[g];
        n1 = nza + za; n2 = nzb + zb;
    } else {
        za = gr0 > nza ? gr0 - nza : 0u;
        zb = grg > nzb ? grg - nzb : 0u;
        n1 = gr0; n2 = grg;
    }
    TargetCtx<K> t;
    t.ref = ref;
    t.mode = J.mode;
    t.vmin = st.vmin;
    t.za = za;
    t.S = 0;
    t.TT = 0;
    K* sp = tgt + start;
    const uint32_t nch = (nzb + kChunk - 1) / kChunk;
    for (uint32_t q = 0; q < nch; ++q) {
        const uint32_t m = min((uint32_t)kChunk, nzb - q * kChunk);
        const bool store = nch > 1;
        if (m <= 64) target_chunk<V, 1>(t, sp, m, q, store);
        else if (m <= 128) target_chunk<V, 2>(t, sp, m, q, store);
        else if (m <= 256) target_chunk<V, 4>(t, sp, m, q, store);
        else if (m <= 512) target_chunk<V, 8>(t, sp, m, q, store);
        else target_chunk<V, 16>(t, sp, m, q, store);
    }
    const uint64_t S = wave_sum_u64(t.S);
    const uint64_t TT = wave_sum_u64(t.TT);
    if (lane == 0) {
        uint32_t ltz;
        if (J.mode == kZeroMin) ltz = 0;
        else if (J.mode == kZeroMax) ltz = nza;
        else ltz = st.refneg;
        const uint64_t S2 = S + (uint64_t)zb * (2ull * ltz + za);
        const uint64_t tie = TT + st.TA + tie_f((uint64_t)za + zb);
        const uint64_t twoN = 2ull * (uint64_t)n1 * (uint64_t)n2;
        const uint64_t u1x2 = twoN >= S2 ? twoN - S2 : 0ull;
        const double U1 = (double)u1x2 * 0.5;
        const double tied = J.tie_corr ? (double)tie : 0.0;
        const int64_t idx = c * J.T + (int64_t)(g - 1);
        J.U1[idx] = U1;
        if (J.tie) J.tie[idx] = tied;
        if (J.n1o) J.n1o[idx] = (double)n1;
        if (J.n2o) J.n2o[idx] = (double)n2;
        if (J.fuse_p) J.P[idx] = p_asymptotic(U1, (double)n1, (double)n2, tied, J.cc, J.alt);
    }
}

// Sort reference keys rk[0..P) (P a power of two >= 64, padded with the
// sentinel) in LDS: registers for runs of up to 1024, LDS merge stages above.
template <typename K>
__device__ __forceinline__ void sort_ref(K* rk, uint32_t P) {
    const int tid = threadIdx.x, nt = blockDim.x;
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
            else if (Kk == 8) HPDEX_SORT_SMALL(8)
            else HPDEX_SORT_SMALL(16)
#undef HPDEX_SORT_SMALL
        }
        __syncthreads();
        return;
    }
    const uint32_t nchunk = P / kChunk;
    for (uint32_t ch = wid; ch < nchunk; ch += nw) {
        K v[16];
        K* b = rk + (size_t)ch * kChunk;
        for (int i = 0; i < 16; ++i) v[i] = b[i * kWave + lane];
        wave_sort<16, K>(v, (int)(ch * kChunk));
        for (int i = 0; i < 16; ++i) b[i * kWave + lane] = v[i];
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
            K v[16];
            K* b = rk + (size_t)ch * kChunk;
            for (int i = 0; i < 16; ++i) v[i] = b[i * kWave + lane];
            wave_merge_tail<16, K>(v, (int)k, (int)(ch * kChunk));
            for (int i = 0; i < 16; ++i) b[i * kWave + lane] = v[i];
        }
        __syncthreads();
    }
}

template <typename V>
__device__ __forceinline__ void process_column(const JobDesc& J, int64_t c, unsigned char* smem, int lds,
                               ColState& st) {
    using KT = KeyOf<V>;
    using K = typename KT::K;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int lane = lane_id();
    const int G = J.G;
    const V* vals = (const V*)J.vals;
    const int64_t p0 = J.indptr[c], p1 = J.indptr[c + 1];
    const uint32_t R = (uint32_t)J.R;

    uint32_t* cnt_nz = (uint32_t*)smem;
    uint32_t* cnt_ez = cnt_nz + G;
    uint32_t* seg = cnt_ez + G;
    const int hdr = (int)(((size_t)3 * G * 4 + 15) & ~(size_t)15);
    unsigned char* rb = smem + hdr;
    const int avail = lds - hdr;
    if (avail < 8192) {
        if (tid == 0) raise_status(J, kStatusGroups, c);
        return;
    }

    for (int g = tid; g < G; g += nt) { cnt_nz[g] = 0; cnt_ez[g] = 0; }
    if (tid == 0) {
        st.all_int = 1u; st.vmin = INT_MAX; st.vmax = INT_MIN; st.refneg = 0; st.refcur = 0;
        st.kmin = ~0ull; st.kmax = 0; st.fail = 0; st.TA = 0;
    }
    __syncthreads();

    // ---- A: per-group counts and value statistics ----
    {
        bool li = true;
        int lmin = INT_MAX, lmax = INT_MIN;
        K kmn = KT::kSentinel, kmx = 0;
        uint32_t neg = 0;
        for (int64_t p = p0 + tid; p < p1; p += nt) {
            const uint32_t r = (uint32_t)J.rows[p];
            if (r >= R) continue;
            const uint32_t g = (uint32_t)J.gid[r];
            if (g >= (uint32_t)G) continue;
            const V v = vals[p];
            if (!KT::finite(v)) continue;
            if (v == V(0)) { atomicAdd(&cnt_ez[g], 1u); continue; }
            atomicAdd(&cnt_nz[g], 1u);
            int iv;
            if (KT::small_int(v, iv)) { lmin = min(lmin, iv); lmax = max(lmax, iv); }
            else li = false;
            if (g == 0) {
                const K k = KT::key(v);
                kmn = k < kmn ? k : kmn;
                kmx = k > kmx ? k : kmx;
                neg += (v < V(0)) ? 1u : 0u;
            }
        }
        const bool wall = __all(li ? 1 : 0);
        lmin = wave_min_i(lmin);
        lmax = wave_max_i(lmax);
        kmn = wave_min_k<K>(kmn);
        kmx = wave_max_k<K>(kmx);
        neg = wave_sum_u32(neg);
        if (lane == 0) {
            if (!wall) atomicAnd(&st.all_int, 0u);
            atomicMin(&st.vmin, lmin);
            atomicMax(&st.vmax, lmax);
            atomicAdd(&st.refneg, neg);
            atomicMin((unsigned long long*)&st.kmin, (unsigned long long)kmn);
            atomicMax((unsigned long long*)&st.kmax, (unsigned long long)kmx);
        }
    }
    __syncthreads();

    // ---- errors of this column (mannwhitneyu.cpp:936-1038) ----
    {
        const uint64_t ebase = (uint64_t)c * (uint64_t)(2 * G + 1);
        for (int g = tid; g < G; g += nt) {
            const uint32_t nz = cnt_nz[g], ez = cnt_ez[g], gr = J.group_rows[g];
            if ((uint64_t)nz + ez > gr) atomicMin(J.err_key, (unsigned long long)(ebase + g));
            const uint32_t n = (J.mode == kZeroNone) ? nz + ez : gr;
            if (n < 2) atomicMin(J.err_key, (unsigned long long)(ebase + (g == 0 ? G : G + g)));
        }
    }

    // ---- choose the reference plan ----
    if (tid == 0) {
        const uint32_t na = cnt_nz[0];
        const int64_t range = (st.vmax >= st.vmin) ? (int64_t)st.vmax - st.vmin + 1 : 1;
        uint32_t refb;
        if (J.hist_max > 0 && st.all_int && range <= J.hist_max) {
            st.hist = 1;
            st.nb = (uint32_t)range;
            if (st.vmax < st.vmin) st.vmin = 0;
            refb = 4u * (st.nb + 1);
        } else {
            st.hist = 0;
            uint32_t P = 64;
            while (P < na) P <<= 1;
            uint32_t nbk = 64;
            while (nbk < na && nbk < (uint32_t)kMaxBuckets) nbk <<= 1;
            st.P = P;
            st.nb = nbk;
            const uint64_t rng = (na > 0) ? st.kmax - st.kmin : 0ull;
            const uint32_t bits = rng ? 64u - (uint32_t)__clzll(rng) : 0u;
            const uint32_t lg = 31u - (uint32_t)__clz(nbk);
            st.shift = bits > lg ? bits - lg : 0u;
            refb = (uint32_t)sizeof(K) * P + 4u * (nbk + 1);
        }
        refb = (refb + 15u) & ~15u;
        if ((int64_t)refb + 64 * (int64_t)sizeof(K) > avail) {
            st.fail = 1;
        } else {
            st.tgt_cap = (uint32_t)((avail - (int)refb) / (int)sizeof(K));
        }
        st.scan[19] = refb;
    }
    __syncthreads();
    if (st.fail) {
        if (tid == 0) raise_status(J, kStatusCapacity, c);
        return;
    }
    const uint32_t refb = st.scan[19];
    K* tgt = (K*)(rb + refb);
    uint32_t* cum = (uint32_t*)rb;
    K* rk = (K*)rb;
    uint32_t* bidx = (uint32_t*)(rb + sizeof(K) * (size_t)st.P);
    const uint32_t hist = st.hist;

    // target segment starts: exclusive prefix over targets 1..G-1
    for (int g = tid; g < G; g += nt) seg[g] = (g == 0) ? 0u : cnt_nz[g];
    __syncthreads();
    const uint32_t ttotal = block_excl_scan(seg + 1, G - 1, st.scan);
    (void)ttotal;
    // any single target too large for the LDS region?  (st.fail was read by
    // every thread before the scan's barriers; a second flag avoids a race)
    if (tid == 0) st.ttotal = 0u;
    __syncthreads();
    for (int g = 1 + tid; g < G; g += nt)
        if (cnt_nz[g] > st.tgt_cap) st.ttotal = 1u;
    if (hist) {
        for (uint32_t b = tid; b <= st.nb; b += nt) cum[b] = 0u;
    }
    __syncthreads();
    if (st.ttotal) {
        if (tid == 0) raise_status(J, kStatusCapacity, c);
        return;
    }

    const int32_t vmin = st.vmin;
    uint32_t glo = 1;
    bool first = true;
    while (glo < (uint32_t)G) {
        // batch [glo, ghi): targets whose keys fit from seg[glo] on
        if (tid == 0) {
            const uint32_t off0 = seg[glo];
            uint32_t lo = glo, hi = (uint32_t)G;   // last g with seg[g]+nz[g]-off0 <= cap
            // binary search for the first g > glo whose end exceeds the cap
            uint32_t a = glo + 1, b = (uint32_t)G;
            while (a < b) {
                const uint32_t mid = (a + b) >> 1;
                if (seg[mid] + cnt_nz[mid] - off0 <= st.tgt_cap) a = mid + 1; else b = mid;
            }
            hi = a;
            (void)lo;
            st.glo = glo; st.ghi = hi; st.off0 = off0;
            st.next_t = glo;
        }
        __syncthreads();
        const uint32_t ghi = st.ghi, off0 = st.off0;

        // ---- C: scatter keys (reference once, targets of this batch) ----
        for (int64_t p = p0 + tid; p < p1; p += nt) {
            const uint32_t r = (uint32_t)J.rows[p];
            if (r >= R) continue;
            const uint32_t g = (uint32_t)J.gid[r];
            if (g >= (uint32_t)G) continue;
            const V v = vals[p];
            if (!KT::finite(v) || v == V(0)) continue;
            if (g == 0) {
                if (first) {
                    if (hist) atomicAdd(&cum[(uint32_t)((int)v - vmin)], 1u);
                    else rk[atomicAdd(&st.refcur, 1u)] = KT::key(v);
                }
            } else if (g >= glo && g < ghi) {
                const uint32_t pos = atomicAdd(&seg[g], 1u) - off0;
                tgt[pos] = hist ? (K)(uint32_t)((int)v - vmin) : KT::key(v);
            }
        }
        __syncthreads();

        // ---- B: reference table ----
        if (first) {
            const uint32_t na = cnt_nz[0];
            uint64_t ta = 0;
            if (hist) {
                for (uint32_t b = tid; b < st.nb; b += nt) ta += tie_f(cum[b]);
                ta = block_sum_u64(ta, st.red);
                block_excl_scan(cum, (int)st.nb + 1, st.scan);
            } else {
                const uint32_t P = st.P;
                for (uint32_t x = na + tid; x < P; x += nt) rk[x] = KT::kSentinel;
                __syncthreads();
                sort_ref<K>(rk, P);
                for (uint32_t b = tid; b <= st.nb; b += nt) bidx[b] = 0u;
                __syncthreads();
                const uint64_t kmin = st.kmin;
                const uint32_t sh = st.shift;
                for (uint32_t x = tid; x < na; x += nt)
                    atomicAdd(&bidx[(uint32_t)(((uint64_t)rk[x] - kmin) >> sh)], 1u);
                __syncthreads();
                block_excl_scan(bidx, (int)st.nb + 1, st.scan);
                RefView<K> rv;
                rv.cum = cum; rv.keys = rk; rv.bidx = bidx; rv.kmin = st.kmin; rv.kmax = st.kmax;
                rv.shift = st.shift; rv.na = na; rv.hist = 0;
                for (uint32_t x = tid; x < na; x += nt) {
                    uint32_t lt, eq;
                    rv.lookup(rk[x], lt, eq);
                    if (lt == x) ta += tie_f(eq);
                }
                ta = block_sum_u64(ta, st.red);
            }
            if (tid == 0) st.TA = ta;
            __syncthreads();
            first = false;
        }

        // ---- D: targets of this batch, one wave per target ----
        {
            RefView<K> rv;
            rv.cum = cum; rv.keys = rk; rv.bidx = bidx; rv.kmin = st.kmin; rv.kmax = st.kmax;
            rv.shift = st.shift; rv.na = cnt_nz[0]; rv.hist = hist;
            for (;;) {
                uint32_t g = 0;
                if (lane == 0) g = atomicAdd(&st.next_t, 1u);
                g = (uint32_t)__builtin_amdgcn_readfirstlane((int)shfl_u32(g, 0));
                if (g >= ghi) break;
                process_target<V>(J, c, g, st, cnt_nz, cnt_ez, seg, tgt, rv);
            }
        }
        __syncthreads();
        glo = ghi;
    }
}

template <typename V>
__global__ __launch_bounds__(512) void gene_plain_kernel(JobDesc J, int lds) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ ColState st;
    process_column<V>(J, (int64_t)blockIdx.x, smem, lds, st);
}

template <typename V>
__global__ __launch_bounds__(512) void gene_persistent_kernel(const JobDesc* jobs,
                                                               const WorkItem* items,
                                                               int n_items, unsigned* head,
                                                               unsigned* done, int lds) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ ColState st;
    for (;;) {
        if (threadIdx.x == 0) st.item = (int32_t)atomicAdd(head, 1u);
        __syncthreads();
        const int it = st.item;
        if (it >= n_items) break;
        const WorkItem w = items[it];
        const JobDesc J = jobs[w.job];
        process_column<V>(J, (int64_t)w.col, smem, lds, st);
        __syncthreads();
    }
    // The last workgroup to leave rewinds the queue for the next launch.
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(done, 1u) == gridDim.x - 1) {
            atomicExch(head, 0u);
            atomicExch(done, 0u);
        }
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
    const int tid = threadIdx.x, nt = blockDim.x;
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
template <typename V>
__global__ __launch_bounds__(1024) void group_mean_kernel(const V* vals, const int32_t* rows,
                                                          const int64_t* indptr,
                                                          const int32_t* gid,
                                                          const uint32_t* group_rows,
                                                          int32_t R, int32_t G, int include_zeros,
                                                          double* out) {
    extern __shared__ __align__(16) unsigned char smem[];
    double* sum = (double*)smem;
    uint32_t* vc = (uint32_t*)(sum + G);
    uint32_t* ic = vc + G;
    const int64_t c = blockIdx.x;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int g = tid; g < G; g += nt) { sum[g] = 0.0; vc[g] = 0u; ic[g] = 0u; }
    __syncthreads();
    const int64_t p0 = indptr[c], p1 = indptr[c + 1];
    for (int64_t p = p0 + tid; p < p1; p += nt) {
        const uint32_t r = (uint32_t)rows[p];
        if (r >= (uint32_t)R) continue;
        const uint32_t g = (uint32_t)gid[r];
        if (g >= (uint32_t)G) continue;
        const double v = (double)vals[p];
        if (__builtin_isfinite(v)) { atomicAdd(&sum[g], v); atomicAdd(&vc[g], 1u); }
        else atomicAdd(&ic[g], 1u);
    }
    __syncthreads();
    for (int g = tid; g < G; g += nt) {
        const uint64_t denom = include_zeros ? ((uint64_t)group_rows[g] - (uint64_t)ic[g])
                                             : (uint64_t)vc[g];
        out[c * G + g] = denom > 0 ? sum[g] / (double)denom : 0.0;
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

hipError_t set_gene_kernel_lds_limit(int lds) {
    hipError_t e;
    e = hipFuncSetAttribute((const void*)gene_plain_kernel<float>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)gene_plain_kernel<double>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)gene_persistent_kernel<float>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)gene_persistent_kernel<double>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)group_mean_kernel<float>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute((const void*)group_mean_kernel<double>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, lds);
}

hipError_t launch_gene_plain(const JobDesc& J, bool f64, int block, int lds, hipStream_t s) {
    if (J.C <= 0) return hipSuccess;
    if (f64) gene_plain_kernel<double><<<(unsigned)J.C, block, lds, s>>>(J, lds);
    else gene_plain_kernel<float><<<(unsigned)J.C, block, lds, s>>>(J, lds);
    return hipGetLastError();
}

hipError_t launch_gene_persistent(const JobDesc* jobs, const WorkItem* items, int n_items,
                                  unsigned* head, unsigned* done, bool f64, int grid, int block,
                                  int lds, hipStream_t s) {
    if (n_items <= 0) return hipSuccess;
    if (f64) gene_persistent_kernel<double><<<grid, block, lds, s>>>(jobs, items, n_items, head, done, lds);
    else gene_persistent_kernel<float><<<grid, block, lds, s>>>(jobs, items, n_items, head, done, lds);
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
                             int include_zeros, double* out, uint32_t* status, int block,
                             int lds, hipStream_t s) {
    (void)status;
    if (C <= 0 || G <= 0) return hipSuccess;
    const int need = 16 * G;
    if (need > lds) return hipErrorInvalidValue;
    if (f64)
        group_mean_kernel<double><<<(unsigned)C, block, need, s>>>((const double*)vals, rows, indptr, gid, group_rows, R, G, include_zeros, out);
    else
        group_mean_kernel<float><<<(unsigned)C, block, need, s>>>((const float*)vals, rows, indptr, gid, group_rows, R, G, include_zeros, out);
    return hipGetLastError();
}

}  // namespace hpdex
