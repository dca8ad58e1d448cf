// pool_main.cpp -- stress harness of the host copy pool of hpdex_stage.cpp
// (its CopyPool class is pasted above this file's main by
// tests/test_copy_pool.py): several threads run jobs of random widths at once,
// every part of every job must run exactly once, and the build runs under
// ThreadSanitizer.
int main() {
    std::atomic<long> bad{0};
    std::vector<std::thread> th;
    for (int t = 0; t < 6; ++t)
        th.emplace_back([&, t] {
            std::mt19937 rng(t);
            for (int it = 0; it < 2000; ++it) {
                const int n = 1 + (int)(rng() % 9);
                std::vector<int> hit(n, 0);
                CopyPool::get().run(n, [&](int i) { hit[i] += 1; });
                for (int i = 0; i < n; ++i)
                    if (hit[i] != 1) bad++;
            }
        });
    for (auto& x : th) x.join();
    std::printf("width %d bad %ld\n", CopyPool::get().width(), bad.load());
    return bad != 0;
}
