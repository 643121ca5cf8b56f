// Checks pca_amd/csrc/host_pool.h on the CPU: every index of every run is
// called exactly once, runs of any size back to back, serial and parallel,
// several pools in turn; built with -fsanitize=thread by the CPU test.
#include <cstdio>
#include <random>

#include "host_pool.h"

int main()
{
    std::mt19937 rng(5);
    std::atomic<int> inits{0};
    for (int round = 0; round < 3; round++) {
        pca::HostPool pool(7, [&] { inits++; });
        for (int it = 0; it < 2000; it++) {
            const size_t n = rng() % 40;
            std::vector<std::atomic<int>> hit(n);
            for (auto &h : hit) h = 0;
            pool.run(n, [&](size_t i) { hit[i]++; }, (rng() & 3) != 0);
            for (size_t i = 0; i < n; i++)
                if (hit[i] != 1) {
                    std::printf("FAIL round %d run %d index %zu called %d times\n", round, it, i, (int)hit[i]);
                    return 1;
                }
        }
    }
    if (inits != 21) {
        std::printf("FAIL worker init ran %d times\n", (int)inits);
        return 1;
    }
    std::printf("ok host pool\n");
    return 0;
}
