// cli.cpp -- the three drop-in command lines, built from this one file:
//
//   gpu_hybrid        <task.json> <max_kernel_depth> [max_solutions=32] [split_depth=-1] [batch_size=4096]
//                     (gpu/hybrid.cu:631-836: host DFS prefixes + device suffixes)
//   gpu_enumerate_oep <task.json> <max_kernel_depth> [max_solutions=32]
//                     (gpu/evaluate_oep.cu:680-944: everything on the device,
//                      OEP tables in device memory)
//   gpu_wavefront     <task.json> <max_kernel_depth> [max_solutions=32] [wavefront_depth=1]
//                     (gpu/wavefront.cu:708-1017: device level expansion, then suffixes)
//
// Arguments, stdout lines and exit codes follow those mains (0 = a solution
// was found, 1 = none or error).  Unlike them, every depth is searched to
// the end and the printed solutions are the first max_solutions of the
// exact, canonically ordered set (the reference CPU solver's DFS order),
// each re-verified on the CPU VM.
#include <chrono>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "pcagpu.h"

#ifndef PCA_CLI_MODE
#define PCA_CLI_MODE 0
#endif
enum { MODE_HYBRID = 0, MODE_ENUM_OEP = 1, MODE_WAVEFRONT = 2 };

static double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static void print_solutions(const pca_task_t &task, const std::vector<pca_solution_t> &sols, long stored)
{
    for (long i = 0; i < stored; i++) {
        const pca_solution_t &s = sols[(size_t)i];
        int passed = pcagpu_task_evaluate(&task, s.prog, s.len);
        std::printf("\nSolution #%ld  len=%d  cpu_verify=%d/%d %s\n", i + 1, s.len, passed, task.num_tests,
                    passed == task.num_tests ? "OK" : "FAIL");
        for (int j = 0; j < s.len; j++) {
            char buf[64];
            pcagpu_disasm(s.prog[j], buf, sizeof(buf));
            std::printf("    %02X: %04X  %s\n", j, s.prog[j], buf);
        }
    }
}

int main(int argc, char **argv)
{
    const int mode = PCA_CLI_MODE;
    const char *usage[] = {
        "Usage: gpu_hybrid <task.json> <max_kernel_depth> [max_solutions] [split_depth] [batch_size]\n",
        "Usage: gpu_enumerate_oep <task.json> <max_kernel_depth> [max_solutions]\n",
        "Usage: gpu_wavefront <task.json> <max_kernel_depth> [max_solutions] [wavefront_depth]\n"};
    if (argc < 3) {
        std::fputs(usage[mode], stderr);
        return 1;
    }
    const char *path = argv[1];
    const int max_depth = std::atoi(argv[2]);
    const long max_solutions = argc >= 4 ? std::atol(argv[3]) : 32;
    int user_split = -1, batch = 4096, wf_depth = 1;
    if (mode == MODE_HYBRID) {
        if (argc >= 5) user_split = std::atoi(argv[4]);
        if (argc >= 6) batch = std::atoi(argv[5]);
    } else if (mode == MODE_WAVEFRONT && argc >= 5) {
        wf_depth = std::atoi(argv[4]);
    }
    if (max_depth < 0 || max_depth > PCAGPU_MAX_KDEPTH) {
        std::fprintf(stderr, "max_kernel_depth must be in [0, %d]\n", PCAGPU_MAX_KDEPTH);
        return 1;
    }
    if (max_solutions < 1 || batch < 1) {
        std::fprintf(stderr, "max_solutions and batch_size must be positive\n");
        return 1;
    }
    pca_task_t task;
    if (pcagpu_task_load(path, &task) < 0) {
        std::fprintf(stderr, "Cannot open task file: %s\n", path);
        return 1;
    }
    if (task.num_outputs != 1) {
        std::fprintf(stderr, "multi-output tasks not supported\n");
        return 1;
    }
    if (task.num_inputs > 7) {
        std::fprintf(stderr, "too many inputs\n");
        return 1;
    }
    if (task.num_inputs + max_depth + 2 > PCAGPU_MAX_PROG) {
        std::fprintf(stderr, "program too long\n");
        return 1;
    }
    pcagpu_opts_t opts = {};
    opts.batch_capacity = (uint32_t)batch;
    pcagpu_t *h = nullptr;
    int rc = pcagpu_open(&task, &opts, &h);
    if (rc < 0) {
        std::fprintf(stderr, "pcagpu_open failed: %d\n", rc);
        return 1;
    }
    const char *title[] = {"PCA-16 hybrid CPU-GPU solver (MI355X)", "PCA-16 exhaustive GPU solver with device OEP (MI355X)",
                           "PCA-16 wavefront GPU solver (MI355X)"};
    std::printf("%s\n", title[mode]);
    std::printf("  task: %s (%s)\n", path, task.name);
    std::printf("  gpu: %s\n", pcagpu_device_name(h));
    std::printf("  tests: %d  inputs: %d  outputs: %d\n", task.num_tests, task.num_inputs, task.num_outputs);
    std::printf("  gpu candidates: %d  batch size: %d\n", 3787, batch);

    const double t0 = now_s();
    int found = -1;
    std::vector<pca_solution_t> sols((size_t)max_solutions);
    std::vector<pca_prefix_t> pre;
    for (int D = 0; D <= max_depth; D++) {
        int split;
        if (mode == MODE_HYBRID) {
            split = user_split >= 0 ? user_split : (D > 0 ? D - 1 : 0);
            if (split > D) split = D;
        } else if (mode == MODE_WAVEFRONT) {
            split = D < wf_depth ? D : wf_depth;
        } else {
            split = D > 0 ? D - 1 : 0;   // device OEP through the last level
        }
        std::printf("\n=== depth %d (split: %s %d + GPU %d) ===\n", D, mode == MODE_HYBRID ? "CPU" : "GPU", split,
                    D - split);
        const double ds = now_s();
        uint64_t nodes = 0, pruned = 0;
        long n;
        long total;
        if (mode == MODE_HYBRID) {
            // one DFS walk streams the prefixes into submit; full batches run
            // on the device while the walk goes on
            pcagpu_solve_opts_t so = {};
            so.max_depth = D;
            so.start_depth = D;
            so.split = split;
            so.producer = 1;
            pcagpu_result_t res;
            if ((rc = pcagpu_solve(h, &so, &res, sols.data(), max_solutions)) < 0) {
                std::fprintf(stderr, "depth %d failed: %d\n", D, rc);
                return 1;
            }
            n = (long)res.prefixes[D];
            nodes = res.prefix_nodes[D];
            pruned = res.prefix_pruned[D];
            total = res.optimal_depth == D ? res.num_solutions : 0;
        } else {
            n = pcagpu_frontier(h, D, split, nullptr, 0);
            if (n < 0) { std::fprintf(stderr, "wavefront failed: %ld\n", n); return 1; }
            pre.resize((size_t)(n > 0 ? n : 1));
            if (n > 0 && pcagpu_frontier(h, D, split, pre.data(), n) < 0) return 1;
            if ((rc = pcagpu_begin_depth(h, D, split)) < 0) return 1;
            for (long i = 0; i < n; i += batch) {
                const uint32_t c = (uint32_t)((n - i) < batch ? (n - i) : batch);
                if ((rc = pcagpu_submit(h, pre.data() + i, c, (uint64_t)i)) < 0) {
                    std::fprintf(stderr, "submit failed: %d\n", rc);
                    return 1;
                }
            }
            total = pcagpu_end_depth(h, sols.data(), max_solutions);
            if (total < 0) {
                std::fprintf(stderr, "depth %d failed: %ld\n", D, total);
                return 1;
            }
        }
        pcagpu_stats_t st;
        pcagpu_get_stats(h, &st);
        if (mode == MODE_HYBRID)
            std::printf("  cpu: %" PRIu64 " nodes, %ld prefixes emitted, OEP pruned: %" PRIu64 ", OEP saturated: 0\n",
                        nodes, n, pruned);
        else
            std::printf("  wavefront: %ld unique prefixes at depth %d\n", n, split);
        std::printf("  elapsed: %.3f sec\n", now_s() - ds);
        if (total > 0) {
            found = D;
            std::printf("\nOptimal solution length: %d instructions (kernel depth %d)\n", task.num_inputs + D + 2, D);
            std::printf("  solutions: %ld (printing the first %ld in canonical order)\n", total,
                        total < max_solutions ? total : max_solutions);
            print_solutions(task, sols, total < max_solutions ? total : max_solutions);
            break;
        }
        std::printf("  solutions: 0\n");
    }
    if (found < 0) std::printf("\nNo solutions found up to kernel depth %d.\n", max_depth);
    pcagpu_stats_t st;
    pcagpu_get_stats(h, &st);
    std::printf("\nTotal: %" PRIu64 " candidates, %.3f sec\n", st.candidates, now_s() - t0);
    std::printf("\nTotal elapsed: %.3f sec\n", now_s() - t0);
    pcagpu_close(h);
    return found >= 0 ? 0 : 1;
}
