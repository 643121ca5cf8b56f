// Checks pca_amd/csrc/host_pool.h on the CPU: every index of every run is
// called exactly once, runs of any size back to back, serial and parallel,
// several pools in turn; every background job runs exactly once, wait_all
// returns only after the jobs submitted before it, jobs run while their
// submitter keeps working, and a queue destroyed with jobs pending runs them
// first; built with -fsanitize=thread by the CPU test.
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
    // background jobs
    for (int round = 0; round < 3; round++) {
        std::atomic<int> jinit{0};
        std::vector<std::atomic<int>> ran(3000);
        for (auto &r : ran) r = 0;
        {
            pca::JobQueue q(round == 0 ? 0u : 3u, [&] { jinit++; });
            size_t next = 0;
            for (int batch = 0; batch < 30; batch++) {
                const size_t n = rng() % 100;
                const size_t first = next;
                for (size_t i = 0; i < n && next < ran.size(); i++, next++) {
                    const size_t id = next;
                    q.submit([&ran, id] { ran[id]++; });
                }
                if (batch % 3 == 0) {
                    q.wait_all();
                    for (size_t i = first; i < next; i++)
                        if (ran[i] != 1) {
                            std::printf("FAIL job %zu ran %d times before wait_all returned\n", i, (int)ran[i]);
                            return 1;
                        }
                }
            }
            // the destructor runs what is still queued
            for (size_t i = 0; i < 50 && next < ran.size(); i++, next++) {
                const size_t id = next;
                q.submit([&ran, id] { ran[id]++; });
            }
            for (size_t i = next; i < ran.size(); i++) ran[i] = 1;   // never submitted
        }
        for (size_t i = 0; i < ran.size(); i++)
            if (ran[i] != 1) {
                std::printf("FAIL round %d job %zu ran %d times\n", round, i, (int)ran[i]);
                return 1;
            }
        if (jinit != (round == 0 ? 0 : 3)) {
            std::printf("FAIL job worker init ran %d times\n", (int)jinit);
            return 1;
        }
    }
    std::printf("ok host pool\n");
    return 0;
}
