// sim_main.cpp -- run the per-gene kernels of hpdex_kernels.hip on the host
// (see sim_runtime.h).  Exposes a C entry point that mirrors the plain
// one-workgroup-per-column launch, for CPU tests and fault hunting.  Built
// with AddressSanitizer by build_sim.py; never shipped.
#include <algorithm>
#include <chrono>
#include <string>
#include <thread>

#include "sim_runtime.h"

thread_local SimDim threadIdx;
SimDim blockIdx, blockDim, gridDim;
namespace sim {
std::vector<std::unique_ptr<Wave>> g_waves;
std::unique_ptr<std::barrier<>> g_block_bar;
unsigned char* g_lds = nullptr;
size_t g_lds_bytes = 0;
thread_local const char* t_last_op = "";
alignas(64) unsigned char g_static_lds[4096];
}  // namespace sim

#include "hpdex_kernels_device.inc"

using namespace hpdex;

namespace {

int sim_last_blocks = 0;   // count-kernel workgroups of the last sim_mwu call
int sim_shrink_range = 0;  // test hook: record each column's range one value short, so the
                           // count kernel meets values outside it (an error status)
int sim_pack = 1;          // the count kernel reads packed entries where a column has them
int sim_split_arrival = 1; // the split keeps entry groups in LDS (tuning split_arrival)
uint32_t sim_ref_cap = 8192; // the LDS/global sort cutover (tuning ref_lds_keys)
int sim_members = 1;       // the member list for columns in row order (job_setup's)
int sim_global_split = 0;  // test hook: the split with its group counters in memory
                           // (split_global_kernel) whatever the group count
int sim_refcache = 0;      // run the job, keep its reference tables (refcache_fill_kernel),
                           // then run it again ranking from the kept tables
int sim_last_cached = 0;   // columns the second run took from the kept tables
std::vector<int32_t> sim_gen_cols;   // sort-path columns of the last run_pipeline
int sim_last_packed = 0;   // count-kernel workgroups of the last call that read packed entries

struct ThreadInfo { std::atomic<const char*>* last; };

// Run one workgroup: nt host threads execute `body` with threadIdx set.
// A watchdog aborts with a per-wave summary if the block does not finish.
bool run_block(unsigned b, unsigned nt, unsigned grid, const std::function<void()>& body,
               double timeout_s) {
    blockIdx.x = b;
    blockDim.x = nt;
    gridDim.x = grid;
    sim::g_waves.clear();
    for (unsigned w = 0; w < nt / 64; ++w) sim::g_waves.emplace_back(new sim::Wave());
    sim::g_block_bar.reset(new std::barrier<>((std::ptrdiff_t)nt));
    // device LDS is not cleared between workgroups: start from garbage
    std::memset(sim::g_lds, 0xA5, sim::g_lds_bytes);
    std::memset(sim::g_static_lds, 0x5A, sizeof(sim::g_static_lds));
    std::vector<std::atomic<const char*>> last(nt);
    std::atomic<unsigned> finished{0};
    std::vector<std::thread> th;
    th.reserve(nt);
    for (unsigned t = 0; t < nt; ++t) {
        th.emplace_back([&, t] {
            threadIdx.x = t;
            body();
            last[t].store("done");
            finished.fetch_add(1);
        });
    }
    const auto t0 = std::chrono::steady_clock::now();
    while (finished.load() < nt) {
        std::this_thread::sleep_for(std::chrono::milliseconds(5));
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (el > timeout_s) {
            std::fprintf(stderr, "sim: block %u stalled after %.1f s; %u of %u threads finished\n",
                         b, el, finished.load(), nt);
            std::fflush(stderr);
            std::_Exit(3);
        }
    }
    for (auto& x : th) x.join();
    return true;
}

}  // namespace

// The launch sequence of launch_mwu_pipeline: column classification, the
// count kernel over the integer-valued columns, then the split / reference /
// target kernels over the other columns.
// Returns the number of columns routed to the count kernel.
template <typename V>
static int run_pipeline(JobDesc& J, int32_t* range, const std::vector<int32_t>& order,
                        uint32_t count_lds, unsigned count_block, double timeout_s) {
    const JobDesc* jobs = &J;
    const int C = (int)J.C;
    std::vector<unsigned char> buf(std::max<size_t>((size_t)kSplitLdsMax, count_lds));
    sim::g_lds = buf.data();
    sim::g_lds_bytes = buf.size();
    std::vector<uint8_t> col_int(C, 0), col_ident(C, 0);
    for (int c = 0; c < C; ++c)
        run_block(c, 256, C, [&] { classify_columns_kernel<V>((const V*)J.vals, J.indptr,
                                                              col_int.data(), range, J.rows,
                                                              (uint32_t)J.R, col_ident.data()); },
                  timeout_s);
    J.colident = col_ident.data();
    J.colint = col_int.data();
    // the matrix's packed entries (finish_matrix), made from the recorded ranges
    const uint32_t vb = 32u - pack_row_bits((uint32_t)J.R);
    std::vector<int32_t> lo(C, INT_MIN);
    std::vector<uint8_t> col_packed(C, 0);
    for (int c = 0; c < C; ++c) {
        const int64_t bins = (int64_t)range[2 * c + 1] - range[2 * c] + 1;
        if (col_int[c] && pack_fits(vb, bins)) {
            lo[c] = bins > 0 ? range[2 * c] : 0;
            col_packed[c] = 1;
        }
    }
    std::vector<uint32_t> packed((size_t)J.indptr[C] + 1, 0xDEADBEEFu);
    std::vector<uint32_t> dense(C, 0), zneg(C, 0);
    for (int c = 0; c < C; ++c)
        run_block(c, 256, C, [&] { pack_columns_kernel<V>((const V*)J.vals, J.rows, J.indptr,
                                                          lo.data(), range, (uint32_t)J.R, vb,
                                                          packed.data(), dense.data(),
                                                          zneg.data()); }, timeout_s);
    J.packed = packed.data();
    J.pack_vb = vb;
    // test hook: ranges one value short of the data, so that the count kernel
    // meets values outside them and hands the columns over
    if (sim_shrink_range)
        for (int c = 0; c < C; ++c)
            if (range[2 * c + 1] > range[2 * c]) range[2 * c + 1] -= 1;
    // the host's routing (job_upload_desc): count workgroups per target slice
    const uint32_t T = (uint32_t)J.T;
    const bool count_ok = count_lds > 0 && J.G >= 2 && (uint32_t)J.G <= kCountMaxGroups;
    std::vector<CountItem> items;
    std::vector<WorkItem> gen_items(C + 1);
    uint32_t n_static = 0;
    unsigned nc = 0;   // counted columns
    for (int32_t c : order) {
        // a packed column's bins are its distinct values; any other's its range
        const bool pk = sim_pack && col_packed[c];
        const int64_t bins = pk ? (int64_t)dense[c] : (int64_t)range[2 * c + 1] - range[2 * c] + 1;
        const int64_t len = J.indptr[c + 1] - J.indptr[c];
        const uint32_t S = (count_ok && col_int[c] && bins <= J.hist_max && len < ((int64_t)1 << 28))
                               ? count_slice_size(T, (uint32_t)(bins > 0 ? bins : 1), len, count_lds)
                               : 0u;
        if (S == 0u) {
            gen_items[n_static++] = WorkItem{0, c};
            continue;
        }
        CountItem it{};
        it.p0 = J.indptr[c];
        it.n = (uint32_t)len;
        it.vmin = pk ? (int32_t)zneg[c] : (bins > 0 ? range[2 * c] : 0);
        it.nb = bins > 0 ? (uint32_t)bins : 1u;
        it.col = c;
        it.S = (uint16_t)S;
        it.flags = pk ? kCountPacked : 0u;
        for (uint32_t t0 = 0; t0 < T; t0 += S) {
            it.t0 = (uint16_t)t0;
            items.push_back(it);
        }
        ++nc;
    }
    sim_last_blocks = (int)items.size();
    sim_last_packed = 0;
    for (const CountItem& x : items) sim_last_packed += (x.flags & kCountPacked) ? 1 : 0;
    const unsigned nb = (unsigned)items.size();
    for (unsigned b = 0; b < nb; ++b)
        run_block(b, count_block, nb, [&] {
            if (count_block == 256)
                count_kernel<V, 256>(J, items.data(), count_lds);
            else if (count_block == 1024)
                count_kernel<V, 1024>(J, items.data(), count_lds);
            else
                count_kernel<V, 512>(J, items.data(), count_lds);
        }, timeout_s);
    const WorkItem* it = gen_items.data();
    sim_gen_cols.clear();
    for (uint32_t k = 0; k < n_static; ++k) sim_gen_cols.push_back(gen_items[k].col);
    // the launch's LDS sizes (launch_mwu_pipeline): the split kernel's most,
    // the reference and target kernels' table
    const uint32_t split_lds = (uint32_t)kSplitLdsMax;
    const bool global_split =
        sim_global_split || split_counter_bytes((uint64_t)J.G) > (uint64_t)split_lds;
    const unsigned ng = n_static;
    for (unsigned b = 0; b < ng; ++b) {
        if (global_split)
            run_block(b, kColBlock, ng, [&] { split_global_kernel<V>(jobs, it, ng); }, timeout_s);
        else
            run_block(b, kSplitBlock, ng, [&] { split_kernel<V>(jobs, it, ng, split_lds); },
                      timeout_s);
    }
    for (unsigned b = 0; b < ng; ++b)
        run_block(b, kRefBlock, ng, [&] { ref_kernel<V>(jobs, it, ng); }, timeout_s);
    if constexpr (sizeof(V) == 4)
        for (unsigned b = 0; b < ng; ++b)
            run_block(b, kTargetBlock, ng, [&] { target_window_kernel<V>(jobs, it, ng); },
                      timeout_s);
    for (unsigned b = 0; b < ng; ++b)
        run_block(b, kColBlock, ng, [&] { target_kernel<V>(jobs, it, ng); }, timeout_s);
    return (int)nc;
}

// count_lds: dynamic LDS of the count kernel (0 = every column takes the
// split/reference/target kernels); block: its workgroup size (256, 512, 1024).  Columns are visited in reversed order.
// Returns the number of columns routed to the count kernel.
extern "C" int sim_mwu(int f64, const void* vals, const int32_t* rows, const int64_t* indptr,
                       int R, int C, const int32_t* gid, int T, int mode, int alt, int tie_corr,
                       int cc, int count_lds, int block, int hist_max, double* U1, double* P,
                       double* tie, double* n1, double* n2, unsigned long long* err_key,
                       unsigned* status, double timeout_s) {
    const unsigned count_block = (block == 256 || block == 1024) ? (unsigned)block : 512u;
    const int G = T + 1;
    std::vector<uint32_t> grows(G, 0u);
    for (int r = 0; r < R; ++r)
        if (gid[r] >= 0 && gid[r] < G) ++grows[gid[r]];
    unsigned long long status_col = ~0ull;
    *err_key = ~0ull;
    *status = 0u;
    const size_t nnz = (size_t)indptr[C];
    std::vector<uint64_t> keys(nnz + 1);
    std::vector<uint32_t> cnt((size_t)C * G), zcnt((size_t)C * G), goff((size_t)C * G);
    std::vector<ColPlan> plan(C);
    std::vector<uint32_t> reftab((size_t)C * kTabStride);
    JobDesc J{};
    const uint32_t bits = label_bits((uint32_t)G);
    std::vector<uint32_t> gidp(label_words((size_t)R, bits));
    pack_labels(gid, (size_t)R, (uint32_t)G, bits, gidp.data());
    // the rows in a group, ascending, with their groups (job_setup builds
    // them for matrices with columns in row order; here always: the split
    // uses them for those columns only)
    std::vector<uint32_t> members, mgroup;
    for (int r = 0; r < R; ++r)
        if (gid[r] >= 0 && gid[r] < G) {
            members.push_back((uint32_t)r);
            mgroup.push_back((uint32_t)gid[r]);
        }
    members.push_back(0u);
    mgroup.push_back(0u);
    J.members = sim_members ? members.data() : nullptr;
    J.member_group = sim_members ? mgroup.data() : nullptr;
    J.n_members = sim_members ? (uint32_t)(members.size() - 1) : 0u;
    J.vals = vals; J.rows = rows; J.indptr = indptr; J.gid = gid; J.group_rows = grows.data();
    J.gidp = gidp.data();
    J.gid_bits = bits;
    J.gid_words = (uint32_t)gidp.size();
    J.keys = keys.data(); J.cnt = cnt.data(); J.zcnt = zcnt.data(); J.goff = goff.data();
    J.plan = plan.data(); J.reftab = reftab.data();
    J.U1 = U1; J.P = P; J.tie = tie; J.n1o = n1; J.n2o = n2;
    J.err_key = err_key; J.status = status; J.status_col = &status_col;
    J.C = C; J.R = R; J.G = G; J.T = T; J.mode = mode; J.alt = alt; J.tie_corr = tie_corr;
    J.asym = 1; J.hist_max = hist_max; J.cc = cc ? 0.5 : 0.0; J.min_n = 2u;
    J.split_arrival = sim_split_arrival;
    J.ref_cap = std::min<uint32_t>(f64 ? (uint32_t)kRefKeysF64 : (uint32_t)kRefKeysF32,
                                   sim_ref_cap);
    // spill space of the columns whose reference slice may pass the on-chip
    // sort (job_upload_desc), for every column: any may take the sort path
    const size_t cap = J.ref_cap;
    std::vector<int64_t> spill_off(C, -1);
    size_t spill_n = 0;
    for (int c = 0; c < C; ++c) {
        const size_t b = std::min((size_t)(indptr[c + 1] - indptr[c]), (size_t)grows[0]);
        if (b > cap) { spill_off[c] = (int64_t)spill_n; spill_n += b; }
    }
    std::vector<uint64_t> spill(spill_n + 1, 0xA5A5A5A5A5A5A5A5ull);
    J.spill = spill.data();
    J.spill_off = spill_off.data();
    std::vector<int32_t> range(2 * (size_t)C + 2);
    J.colrange = range.data();
    std::vector<int32_t> order(C);
    for (int c = 0; c < C; ++c) order[c] = C - 1 - c;   // reversed order
    auto run = [&] {
        if (f64)
            return run_pipeline<double>(J, range.data(), order, (uint32_t)count_lds,
                                        count_block, timeout_s);
        return run_pipeline<float>(J, range.data(), order, (uint32_t)count_lds, count_block,
                                   timeout_s);
    };
    sim_last_cached = 0;
    if (!sim_refcache) return run();
    // the host's refcache_fill: offsets from the column plans, then the fill
    // kernel copies the sorted keys and bucket indexes
    run();
    if (*err_key != ~0ull || *status != 0u) return -1;
    std::vector<RefCacheCol> rc(C);
    std::memset(rc.data(), 0, rc.size() * sizeof(RefCacheCol));
    std::vector<int32_t> list;
    int64_t koff = 0, toff = 0;
    for (int32_t c : sim_gen_cols) {
        if (plan[c].hist) continue;
        rc[c].koff = koff; rc[c].toff = toff; rc[c].TA = plan[c].TA;
        rc[c].kmin = plan[c].kmin; rc[c].kmax = plan[c].kmax;
        rc[c].shift = plan[c].shift; rc[c].refneg = plan[c].refneg;
        rc[c].na = plan[c].na; rc[c].nb = plan[c].nb;
        koff += plan[c].na;
        toff += (int64_t)plan[c].nb + 1;
        list.push_back(c);
    }
    std::vector<uint64_t> rkeys((size_t)koff + 1);
    std::vector<uint32_t> rtab((size_t)toff + 1);
    const unsigned nl = (unsigned)list.size();
    std::vector<unsigned char> lds(4096);   // the fill kernel takes no dynamic LDS
    sim::g_lds = lds.data();
    sim::g_lds_bytes = lds.size();
    for (unsigned b = 0; b < nl; ++b)
        run_block(b, kColBlock, nl, [&] {
            if (f64) refcache_fill_kernel<uint64_t>(&J, list.data(), nl, rc.data(),
                                                    (uint64_t*)rkeys.data(), rtab.data());
            else refcache_fill_kernel<uint32_t>(&J, list.data(), nl, rc.data(),
                                                (uint32_t*)rkeys.data(), rtab.data());
        }, timeout_s);
    // the job's own tables and plans are garbage now: the second run must
    // take every kept column's reference from the cache
    std::memset(reftab.data(), 0xAB, reftab.size() * 4);
    std::memset(plan.data(), 0xCD, plan.size() * sizeof(ColPlan));
    std::memset(keys.data(), 0xEF, keys.size() * 8);
    J.rc = rc.data();
    J.rc_keys = rkeys.data();
    J.rc_tab = rtab.data();
    *err_key = ~0ull;
    *status = 0u;
    status_col = ~0ull;
    const int r = run();
    sim_last_cached = (int)nl;
    return r;
}

// Count-kernel workgroups (column x target slice) of the last sim_mwu call.
extern "C" int sim_count_blocks() { return sim_last_blocks; }
extern "C" void sim_set_shrink_range(int on) { sim_shrink_range = on; }
extern "C" void sim_set_pack(int on) { sim_pack = on; }
extern "C" int sim_packed_blocks() { return sim_last_packed; }
extern "C" void sim_set_split_arrival(int on) { sim_split_arrival = on; }
extern "C" void sim_set_ref_cap(int keys) { sim_ref_cap = (uint32_t)keys; }
extern "C" void sim_set_refcache(int on) { sim_refcache = on; }
extern "C" void sim_set_global_split(int on) { sim_global_split = on; }
extern "C" void sim_set_members(int on) { sim_members = on; }
extern "C" int sim_cached_columns() { return sim_last_cached; }
