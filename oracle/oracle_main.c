/*
 * oracle_main.c -- command line of the CPU oracle (test infrastructure).
 *
 *   oracle_enumerate <task.json> [-d D] [--no-branch] [--no-oep] [-v]
 *                    [--exact | --ref32] [--max-solutions N]
 *                    [--start-depth N] [--time-limit S] [--json]
 *
 * Argument handling and stdout follow tools/enumerate.c:537-662 so that the
 * default (--ref32) run is byte-comparable with the reference binary apart
 * from wall-clock figures.  --json prints one machine-readable line instead.
 */
#include "pca_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr,
                "Usage: oracle_enumerate <task.json> [-d max_depth] [--no-branch] [--no-oep]\n"
                "       [--exact|--ref32] [--max-solutions N] [--start-depth N]\n"
                "       [--time-limit S] [--json]\n");
        return 1;
    }
    or_opts_t o;
    or_opts_default(&o);
    o.oep_mode = OR_OEP_REF32;
    o.shadow_exact = 1;
    o.print = 1;
    int json = 0;
    for (int i = 2; i < argc; i++) {
        if (!strcmp(argv[i], "-d") && i + 1 < argc) o.max_depth = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--no-branch")) o.use_branches = 0;
        else if (!strcmp(argv[i], "--no-oep")) o.oep_mode = OR_OEP_OFF;
        else if (!strcmp(argv[i], "--exact")) o.oep_mode = OR_OEP_EXACT;
        else if (!strcmp(argv[i], "--ref32")) o.oep_mode = OR_OEP_REF32;
        else if (!strcmp(argv[i], "--no-shadow")) o.shadow_exact = 0;
        else if (!strcmp(argv[i], "--max-solutions") && i + 1 < argc) o.max_solutions = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--start-depth") && i + 1 < argc) o.start_depth = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--time-limit") && i + 1 < argc) o.time_limit = atof(argv[++i]);
        else if (!strcmp(argv[i], "--json")) { json = 1; o.print = 0; }
        else if (!strcmp(argv[i], "-v")) { /* verbose: no extra output in pass 1 */ }
    }
    or_task_t task;
    if (or_task_load(argv[1], &task) < 0) return 1;
    or_result_t res;
    or_solve(&task, &o, &res);
    if (json) {
        printf("{\"optimal_depth\": %d, \"num_solutions\": %d, \"timed_out\": %d, "
               "\"seconds\": %.6f, \"cap_overflow\": %d, \"tried\": [",
               res.optimal_depth, res.num_solutions, res.timed_out, res.seconds, res.cap_overflow);
        for (int d = 0; d <= o.max_depth && d < 17; d++) printf("%s%llu", d ? ", " : "", (unsigned long long)res.tried[d]);
        printf("], \"leaf_tried\": [");
        for (int d = 0; d <= o.max_depth && d < 17; d++) printf("%s%llu", d ? ", " : "", (unsigned long long)res.leaf_tried[d]);
        printf("], \"pruned\": [");
        for (int d = 0; d <= o.max_depth && d < 17; d++) printf("%s%llu", d ? ", " : "", (unsigned long long)res.pruned[d]);
        printf("], \"ref_events\": [");
        for (int d = 0; d <= o.max_depth && d < 17; d++)
            printf("%s%llu", d ? ", " : "", (unsigned long long)(res.ref_collisions[d] + res.ref_saturations[d]));
        printf("], \"solutions\": [");
        for (int i = 0; i < res.num_solutions; i++) {
            printf("%s[", i ? ", " : "");
            for (int j = 0; j < res.solution_lens[i]; j++) printf("%s%u", j ? ", " : "", res.solutions[i][j]);
            printf("]");
        }
        printf("]}\n");
    } else {
        long ev = 0;
        for (int d = 0; d < 17; d++) ev += (long)(res.ref_collisions[d] + res.ref_saturations[d]);
        if (o.oep_mode == OR_OEP_REF32 && o.shadow_exact)
            fprintf(stderr, "oracle: ref32 divergence events (collision prunes + saturations): %ld\n", ev);
    }
    int rc = res.num_solutions > 0 ? 0 : 1;
    or_result_free(&res);
    return rc;
}
