// compact_main.cpp -- harness of the compact upload's host fill of
// hpdex_stage.cpp (its compact_run / compact_fill code is pasted above this
// file's main by tests/test_copy_pool.py).  Random CSC matrices are filled in
// random chunks, decoded the way the device decodes them (a column's high row
// bits start at its first row's and go up by one wherever the low bits fall)
// and compared with the source; entries that must not qualify are planted one
// at a time and the chunk holding them must report their column, and only it.
#include <cstring>

template <typename IDX>
static int run_case(std::mt19937_64& g, int64_t R, int C, double dens) {
    std::vector<int64_t> indptr(C + 1, 0);
    std::vector<IDX> rows;
    std::vector<float> vals;
    for (int c = 0; c < C; ++c) {
        const double d = (c % 5 == 0) ? 0.0 : dens * (1 + (int)(g() % 3));
        for (int64_t r = 0; r < R; ++r)
            if ((double)(g() % 1000000) < d * 1e6) {
                rows.push_back((IDX)r);
                vals.push_back((float)(g() % 7 == 0 ? 0 : g() % 66000 % 65536));
            }
        indptr[c + 1] = (int64_t)rows.size();
    }
    const size_t nnz = rows.size();
    std::vector<uint32_t> out(nnz + 1);
    // random chunk boundaries
    size_t off = 0;
    while (off < nnz) {
        const size_t m = std::min(nnz - off, (size_t)(1 + g() % 5000));
        int nbad = 0;
        compact_fill<IDX>(out.data() + off, vals.data(), rows.data(), indptr.data(), (size_t)C,
                          off, m, R, [&](size_t, size_t) { ++nbad; });
        if (nbad) {
            std::printf("clean matrix refused at %zu\n", off);
            return 1;
        }
        off += m;
    }
    // decode as the device does
    for (int c = 0; c < C; ++c) {
        if (indptr[c] == indptr[c + 1]) continue;
        uint32_t hi = (uint32_t)((int64_t)rows[indptr[c]] >> 16), prev = 0xFFFFFFFFu;
        for (int64_t p = indptr[c]; p < indptr[c + 1]; ++p) {
            const uint32_t lo = out[p] >> 16;
            if (prev != 0xFFFFFFFFu && lo < prev) ++hi;
            prev = lo;
            const int64_t r = ((int64_t)hi << 16) | lo;
            const float v = (float)(out[p] & 0xFFFFu);
            if (r != (int64_t)rows[p] || std::memcmp(&v, &vals[p], 4) != 0) {
                std::printf("decode differs at %ld\n", (long)p);
                return 1;
            }
        }
    }
    // every kind of entry that must not qualify, at random places
    int bad = 0;
    for (int kind = 0; kind < 9 && nnz > 4; ++kind) {
        for (int rep = 0; rep < 20; ++rep) {
            std::vector<IDX> r2 = rows;
            std::vector<float> v2 = vals;
            const size_t at = 1 + g() % (nnz - 2);
            const int c = (int)(std::upper_bound(indptr.begin(), indptr.end(), (int64_t)at) -
                                indptr.begin()) - 1;
            const bool has_prev = (int64_t)at > indptr[c];
            switch (kind) {
            case 0: v2[at] = 2.5f; break;
            case 1: v2[at] = 65536.0f; break;
            case 2: v2[at] = -3.0f; break;
            case 3: v2[at] = -0.0f; break;
            case 4: v2[at] = std::numeric_limits<float>::quiet_NaN(); break;
            case 5: v2[at] = std::numeric_limits<float>::infinity(); break;
            case 6: r2[at] = (IDX)R; break;                       // out of range
            case 7: if (!has_prev) continue; r2[at] = r2[at - 1]; break;   // repeated
            case 8: if (!has_prev) continue; r2[at] = r2[at - 1] + (IDX)65536; break;  // gap 2^16
            }
            if (kind == 8 && (int64_t)r2[at] >= R) continue;
            // the chunk holding `at` (one that starts before it and one at it)
            const size_t a = at - std::min(at, (size_t)(g() % 300));
            const size_t m = std::min(nnz - a, (size_t)(1 + (at - a) + g() % 300));
            std::vector<size_t> cols;
            compact_fill<IDX>(out.data(), v2.data(), r2.data(), indptr.data(), (size_t)C, a, m,
                              R, [&](size_t col, size_t) { cols.push_back(col); });
            if (cols.size() != 1 || cols[0] != (size_t)c) {
                std::printf("kind %d at %zu (column %d): %zu columns reported\n", kind, at, c,
                            cols.size());
                ++bad;
            }
        }
    }
    return bad;
}

int main() {
    std::mt19937_64 g(7);
    int bad = 0;
    bad += run_case<int32_t>(g, 3000, 40, 0.05);
    bad += run_case<int64_t>(g, 3000, 40, 0.05);
    bad += run_case<int32_t>(g, 300000, 12, 0.01);   // low row bits wrap
    bad += run_case<int64_t>(g, 300000, 12, 0.002);  // gaps up to a few thousand rows
    std::printf("bad %d\n", bad);
    return bad != 0;
}
