// The edit path's host work pool: four host threads submit 200 jobs each;
// every index is written exactly by its own job (built under ThreadSanitizer).
#include "vgi_pool.h"
#include <cstdio>
#include <numeric>
int main() {
    vgi_host::WorkPool pool(7);
    bool ok = true;
    auto job = [&](int seed) {
        for (int rep = 0; rep < 200; ++rep) {
            size_t n = 1000 + (size_t)(seed * 37 + rep * 11) % 5000;
            std::vector<long> out(n, -1);
            pool.run(n, [&](size_t b, size_t e) { for (size_t i = b; i < e; ++i) out[i] = (long)(i * i + seed); });
            for (size_t i = 0; i < n; ++i) if (out[i] != (long)(i * i + seed)) ok = false;
        }
    };
    std::vector<std::thread> callers;
    for (int c = 0; c < 4; ++c) callers.emplace_back(job, c);
    for (auto &c : callers) c.join();
    printf("%s\n", ok ? "pool ok" : "POOL WRONG");
    return ok ? 0 : 1;
}
