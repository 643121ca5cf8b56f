/*
 * pcagpu.h -- C ABI of the MI355X search path for the PCA-16 superoptimizer.
 *
 * The reference has no library interface: each GPU solver is a CLI whose
 * main() owns an in-process host<->device seam.  The entry points below are
 * that seam, cut out as plain C so that a host DFS (C, C++, Python ctypes)
 * can drive hand-written HIP kernels for gfx950.  Plain pointers and sizes
 * only; no HIP or torch types cross this boundary.
 *
 * Conventions: functions return 0 (or a count) on success and a negative
 * PCAGPU_E* code on failure; the library never exits the process.  Handles
 * are independent (no global state), so several tasks can be resident at
 * once; one handle must not be used from two threads at the same time.
 * The caller owns input memory; the library copies what it keeps.
 */
#ifndef PCAGPU_H
#define PCAGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PCAGPU_MAX_TESTS   32   /* tasks/spec.h:17 MAX_TESTS                */
#define PCAGPU_MAX_PORTS   8    /* tasks/spec.h:18 MAX_IO_PORTS             */
#define PCAGPU_MAX_KDEPTH  8    /* gpu/hybrid.cu:29 MAX_KERNEL_DEPTH        */
#define PCAGPU_MAX_PROG    16   /* gpu/hybrid.cu:30 MAX_PROG_LEN            */

/* error codes */
#define PCAGPU_OK              0
#define PCAGPU_EINVAL         -1   /* bad argument / unsupported task        */
#define PCAGPU_ENOMEM         -2   /* device or host allocation failed       */
#define PCAGPU_EHIP           -3   /* a HIP runtime call failed              */
#define PCAGPU_ESTATE         -4   /* call out of order                      */
#define PCAGPU_EOVERFLOW      -5   /* a fixed-capacity buffer overflowed      */
#define PCAGPU_ECOLLISION     -6   /* two distinct OEP states shared a 64-bit
                                      key: the exact path refuses to guess   */
#define PCAGPU_ENODEV         -7   /* no GPU visible                          */
#define PCAGPU_EIO            -8   /* task file unreadable                    */
#define PCAGPU_EVERIFY        -9   /* a device result failed host
                                      re-verification (an internal error)    */

/* Task record, byte-compatible with task_spec_t (tasks/spec.h:20-34). */
typedef struct {
    uint16_t in[PCAGPU_MAX_PORTS];
    uint16_t out[PCAGPU_MAX_PORTS];
} pca_test_case_t;

typedef struct {
    char            name[64];
    uint64_t        max_cycles;
    uint8_t         input_ports[PCAGPU_MAX_PORTS];
    int             num_inputs;
    uint8_t         output_ports[PCAGPU_MAX_PORTS];
    int             num_outputs;
    pca_test_case_t tests[PCAGPU_MAX_TESTS];
    int             num_tests;
} pca_task_t;

/* CPU->GPU unit of work, same 18-byte layout as prefix_t
 * (gpu/hybrid.cu:56-60): the first num_insns kernel instructions and the
 * live-register mask after them (bit i = r_i holds a written value). */
#pragma pack(push, 1)
typedef struct {
    uint16_t insns[PCAGPU_MAX_KDEPTH];
    uint8_t  num_insns;
    uint8_t  live;
} pca_prefix_t;
#pragma pack(pop)

/* A verified program at the target depth, in canonical (reference DFS)
 * order: ordered by (prefix_rank, cand[], out_reg). */
typedef struct {
    uint64_t prefix_rank;            /* DFS rank of its split-depth prefix   */
    uint16_t prog[PCAGPU_MAX_PROG];  /* scaffold + kernel + OUT + HLT        */
    uint8_t  len;                    /* words in prog                        */
    uint8_t  out_reg;                /* register written by OUT              */
    uint8_t  kernel_len;             /* kernel instructions                  */
    uint8_t  cpu_verified;           /* 1: task_evaluate passed every test   */
    uint16_t cand[PCAGPU_MAX_KDEPTH];/* candidate-list index at each kernel
                                        position (the DFS rank path)         */
} pca_solution_t;

typedef struct {
    uint64_t candidates;      /* (prefix, suffix, out_reg) triples that pass
                                 the static grammar/liveness filters         */
    uint64_t suffixes;        /* (prefix, suffix) pairs evaluated            */
    uint64_t tests_evaluated; /* per-candidate test evaluations (early exit) */
    uint64_t prefixes;        /* prefixes submitted                          */
    uint64_t raw_hits;        /* device hits before exact OEP resolution     */
    uint64_t kernel_ns;       /* device time of the scan kernels (HIP events)*/
    uint64_t launches;        /* scan kernel launches                        */
    uint64_t frontier_nodes;  /* wavefront: children generated               */
    uint64_t frontier_kept;   /* wavefront: survivors after exact OEP        */
    uint64_t expand_ns;       /* wavefront kernels' device time              */
    uint64_t oep_collisions;  /* 64-bit key matches with a different state   */
    uint64_t parents;         /* parent states scanned (last-position lists) */
    uint64_t scan_tile;       /* tile of the last scan launch: block_kids |
                                 mask_tests << 8 | mode << 16 (0: last
                                 position shared per block, 1: per child)    */
    uint64_t tests_executed;  /* group tests the scan kernels ran: below
                                 tests_evaluated (the tests deciding every
                                 group, which a per-child scan runs) by what
                                 a block's node decides once for its
                                 children                                    */
} pcagpu_stats_t;

typedef struct {
    int      device;          /* HIP device ordinal (default 0)               */
    uint32_t batch_capacity;  /* prefixes per flush (default 1<<20)           */
    uint32_t hit_capacity;    /* device hit records per depth (default 1<<22) */
    uint32_t waves_per_cu;    /* persistent scan occupancy (0 = autotuned)    */
    uint32_t fixed_test_order;/* 1: keep the initial device test order; 0:
                                 reorder tests 2.. after every scan launch by
                                 how often each ends a surviving candidate   */
    uint32_t scan_tile;       /* low 16 bits 0: autotuned; else block_kids
                                 (1..64) | mask_tests (1..6) << 8, fixed;
                                 bits 16..23: 0 decides the last-position
                                 value groups that do not read the register
                                 a child writes once per block, 1 scans
                                 every child on its own (same results)      */
    uint32_t eager_flush;     /* 1: plain stream calls per scan launch; 0: one
                                 replayed HIP graph (copies, clears, kernel,
                                 counter read-back)                          */
    uint32_t test_flags;      /* testing only (0 in use): bit 0 gives the
                                 wavefront's OEP table 64 slots, so it
                                 overflows (PCAGPU_EOVERFLOW); bit 1 keeps 3
                                 bits of its 64-bit keys, so distinct states
                                 share keys (PCAGPU_ECOLLISION)              */
} pcagpu_opts_t;

typedef struct pcagpu pcagpu_t;

/* ---- host helpers (the reference's C API for load/verify/print) ---- */

/* tasks/spec.c:80-184 task_load: same parser and quirks. */
int  pcagpu_task_load(const char *path, pca_task_t *task);
/* Same, from an in-memory JSON string. */
int  pcagpu_task_parse(const char *json, pca_task_t *task);
/* tasks/spec.c:188-215 task_evaluate: number of tests passed. */
int  pcagpu_task_evaluate(const pca_task_t *task, const uint16_t *prog, int len);
/* src/asm.c:420-503 pca_disasm: identical text. */
void pcagpu_disasm(uint16_t insn, char *buf, size_t len);
/* tools/enumerate.c:206-267 gen_candidates (branch mode, use_branches=1):
 * the ordered candidate list at kernel position pos of a depth-D tree. */
int  pcagpu_gen_candidates(uint16_t *out, int max, uint8_t live, int pos, int kernel_depth);
/* Search-state records for the value network (include/pcavalue.h): for
 * each of n kernel prefixes (insns[i*PCAGPU_MAX_KDEPTH ..], len[i] grammar
 * instructions at kernel positions 0..len[i]-1) and each test t < 32, the
 * 16 words rec[(i*32 + t)*16 ..] = {input0, input1, target, r0..r7, flags,
 * frozen (a DIV/MOD fault stopped the VM), resume (pending forward branch
 * target), 0, 0} of the state after the IN scaffold and the prefix (the state
 * the reference serialises, src/search_state.h:38-61); tests >= num_tests
 * are all zero.  live[i] = the prefix's live mask (r1..r_n_in plus every
 * register written).  No counterpart in the reference's C API: its Python
 * feature extractor reads search_state_t (python/search_eval.py:206-251). */
int  pcagpu_state_records(const pca_task_t *task, const uint16_t *insns, const uint8_t *len, int n,
                          uint16_t *rec, uint8_t *live);

/* ---- device handle: replaces the constant uploads and buffers of
 *      gpu/hybrid.cu:592-621, :662-680 ---- */
int  pcagpu_open(const pca_task_t *task, const pcagpu_opts_t *opts, pcagpu_t **out);
void pcagpu_close(pcagpu_t *h);
const char *pcagpu_device_name(pcagpu_t *h);
/* Re-points an open handle at another task: only the task record is
 * uploaded (grammar tables, batch and hit buffers stay), and any depth in
 * progress is dropped.  No reference counterpart (the reference loads one
 * task per process): it lets one resident handle serve tasks one after
 * another, as the server's slots do. */
int  pcagpu_retask(pcagpu_t *h, const pca_task_t *task);

/* Start kernel depth D with the CPU prefixes ending at split s
 * (gpu/hybrid.cu:699-745): resets the depth's hit list. 0 <= s <= D. */
int  pcagpu_begin_depth(pcagpu_t *h, int target_depth, int split_depth);

/* Queue n split-depth prefixes whose DFS ranks are first_rank..+n-1
 * (gpu/hybrid.cu:507-521 emit).  Copies them; flushes automatically when the
 * batch is full.  Ranks must increase across calls within a depth. */
int  pcagpu_submit(pcagpu_t *h, const pca_prefix_t *prefixes, uint32_t n, uint64_t first_rank);

/* Evaluate every queued prefix against every suffix of length D - s
 * (gpu/hybrid.cu:470-505 flush).  Blocks until done.  Returns the number of
 * raw device hits so far in this depth (>= 0) or an error. */
int  pcagpu_flush(pcagpu_t *h);

/* Finish the depth: exact OEP resolution of the raw hits (reference
 * first-visitor semantics with exact keys), canonical ordering, CPU
 * re-verification (gpu/hybrid.cu:784-816).  Writes up to cap solutions and
 * returns the total count (which may exceed cap), or an error. */
long pcagpu_end_depth(pcagpu_t *h, pca_solution_t *out, long cap);

/* The raw device hits of the current depth before OEP resolution: the
 * brute-force suffix semantics of gpu/hybrid.cu:163-275 restricted to the
 * reference grammar.  Each row is (prefix index, flags, match mask, then the
 * candidate-list indices at positions s..D-1), 14 uint16 per row:
 *   row[0..3] prefix index (little-endian u64), row[4] flags (bit 0: cannot
 *   pass: fault or cycle limit), row[5] mask of registers r1..r7 equal to the
 *   target on every test, row[6..] indices (k of them, rest 0).
 * Rows come in canonical order.  Call after pcagpu_flush and before
 * pcagpu_end_depth.  Returns the total row count. */
long pcagpu_raw_hits(pcagpu_t *h, uint16_t *rows, long cap);

/* ---- allocation-free raw-hit scan ----
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the rows pcagpu_raw_hits returns after
 * pcagpu_begin_depth(D, s) + pcagpu_submit(prefixes, n) + pcagpu_flush,
 * computed without a handle and without allocating.  Every device buffer
 * (grammar tables, task record, prefixes, scan blocks, hit records, launch
 * counters) is one fixed array of the library, about 11 MiB per GPU, held
 * until the process ends; the first call on a device (the current one)
 * sets it up.  Runs on the current device's null stream and blocks until
 * done.  At most PCAGPU_NOALLOC_MAX_PREFIXES prefixes, their scan blocks
 * (k = D - s >= 2: one per 64 children of a node one level above the last;
 * a 3-instruction suffix has up to about 2.3e5 per prefix) at most
 * PCAGPU_NOALLOC_MAX_BLOCKS, and at most PCAGPU_NOALLOC_MAX_HITS raw hits:
 * beyond them it returns PCAGPU_EOVERFLOW and writes nothing.  Writes up to
 * cap rows (14 uint16 each, the pcagpu_raw_hits layout, prefix index into
 * `prefixes`) and returns the row count.  Calls that overlap on one GPU
 * race: they share its array. */
#define PCAGPU_NOALLOC_MAX_PREFIXES 16384
#define PCAGPU_NOALLOC_MAX_BLOCKS   (1 << 18)
#define PCAGPU_NOALLOC_MAX_HITS     (1 << 18)
long pcagpu_raw_hits_noalloc(const pca_task_t *task, int target_depth, int split_depth,
                             const pca_prefix_t *prefixes, uint32_t n, uint16_t *rows, long cap);

/* Exhaustive check of the 16-bit ALU (src/vm.c:79-136 semantics): the sum
 * over a in [0, 2^16), b in [b_lo, b_hi) of a 64-bit mix of (a, b, value,
 * flags, fault) for opcode op (ADD..ASR, CMP).  h == NULL computes it on
 * the host, otherwise on h's device.  No reference counterpart: it exists
 * to pin the device arithmetic against the CPU oracle. */
int  pcagpu_alu_digest(pcagpu_t *h, int op, uint32_t b_lo, uint32_t b_hi, uint64_t *out);

int  pcagpu_get_stats(pcagpu_t *h, pcagpu_stats_t *stats);
void pcagpu_reset_stats(pcagpu_t *h);

/* ---- the GPU wavefront producer (gpu/wavefront.cu:330-414 expand) ----
 * Exact-OEP level-synchronous expansion of the depth-D tree down to split
 * s.  Survivors come back in DFS order with their ranks 0..n-1; BR children
 * are exempt from OEP as in tools/enumerate.c:435.  Returns the survivor
 * count (writes up to cap of them) or an error. */
long pcagpu_frontier(pcagpu_t *h, int target_depth, int split_depth,
                     pca_prefix_t *out, long cap);

/* ---- the CPU prefix producer (gpu/hybrid.cu:524-588 hybrid_enumerate) ----
 * Branch-mode DFS of the depth-D tree down to split s with exact OEP keys
 * and no early stop; no GPU needed.  Writes the first cap surviving prefixes
 * in DFS order and stops there when out is given (returns min(total, cap));
 * with out == NULL it walks the whole level and returns the total.  nodes_out / pruned_out (may be NULL) get the
 * prefix levels' tried and OEP-pruned counters. */
long pcagpu_host_frontier(const pca_task_t *task, int target_depth, int split_depth,
                          pca_prefix_t *out, long cap, uint64_t *nodes_out, uint64_t *pruned_out);

/* ---- whole IDDFS solve (the drop-in for the three CLIs' main loops) ---- */
typedef struct {
    int      max_depth;       /* last kernel depth (<= 8)                      */
    int      start_depth;     /* first kernel depth (default 0)                */
    int      split;           /* -1 = auto (D-1), else fixed split            */
    int      producer;        /* 0 = GPU wavefront, 1 = host DFS (hybrid.cu)  */
    int      verbose;         /* print per-depth progress to stdout           */
    double   time_limit;      /* seconds, 0 = none: stop between flushes      */
} pcagpu_solve_opts_t;

typedef struct {
    int      optimal_depth;   /* -1 = none found up to max_depth              */
    long     num_solutions;   /* total at the optimal depth                   */
    int      depths_done;     /* last depth completely searched               */
    int      timed_out;
    double   seconds;
    uint64_t candidates[PCAGPU_MAX_KDEPTH + 1]; /* per depth                  */
    uint64_t prefixes[PCAGPU_MAX_KDEPTH + 1];
    /* per depth, the prefix levels: nodes tried and OEP-pruned (host DFS,
     * tools/enumerate.c counters), or children generated and dropped as
     * duplicates (GPU wavefront) */
    uint64_t prefix_nodes[PCAGPU_MAX_KDEPTH + 1];
    uint64_t prefix_pruned[PCAGPU_MAX_KDEPTH + 1];
    double   depth_seconds[PCAGPU_MAX_KDEPTH + 1];  /* wall time of each depth */
} pcagpu_result_t;

/* Runs IDDFS on an open handle.  Writes up to cap solutions of the optimal
 * depth.  Returns 0 or an error. */
int  pcagpu_solve(pcagpu_t *h, const pcagpu_solve_opts_t *opts, pcagpu_result_t *res,
                  pca_solution_t *out, long cap);

/* ---- exhaustive leaf completion (src/search_state.c:339-453) ----
 * The branchless search state of the reference's learned-search loop and
 * the brute-force completion it calls near the depth budget
 * (python/bootstrap.py:85-113 via python/search_eval.py:171-181). */

/* Byte-compatible with search_state_t (src/search_state.h:19-25, 546 B):
 * per-test registers r0..r7 and flags, the live mask (bit i: r_i is an input
 * or was written), and the kernel instructions executed so far. */
typedef struct {
    uint16_t regs[PCAGPU_MAX_TESTS][8];
    uint8_t  flags[PCAGPU_MAX_TESTS];
    uint8_t  live;
    uint8_t  depth;
} pca_search_state_t;

/* src/search_state.c:34-46 search_state_init_root: inputs in r1..r_n. */
void    pcagpu_state_init_root(const pca_task_t *task, pca_search_state_t *s);
/* src/search_state.c:49-187 search_state_step: 1 with *dst the state after
 * insn on every test, or 0 (dst unspecified) if insn is not one of ADD..ASR,
 * LDI, ADDI or a DIV/MOD divides by zero on some test. */
int     pcagpu_state_step(const pca_task_t *task, const pca_search_state_t *src, uint16_t insn,
                          pca_search_state_t *dst);
/* src/search_state.c:223-239 search_state_leaf_outputs: mask of live
 * registers r1..r7 equal to the target on every test. */
uint8_t pcagpu_state_leaf_outputs(const pca_task_t *task, const pca_search_state_t *s);
/* src/search_state.c:280-330 search_gen_candidates_branchless: the ordered
 * candidate list of a state with live mask `live`. */
int     pcagpu_gen_candidates_branchless(uint16_t *out, int max, uint8_t live);

/* src/search_state.c:418-445 search_exhaustive_complete on h's device (the
 * handle's task is the reference's search_ctx_t): the shortest completion
 * of `start` by at most max_depth - start->depth branchless instructions,
 * and among those the first in the reference's DFS order, with its OEP
 * (first visitor of each register/flag state per depth, live mask ignored)
 * on exact keys.  Returns 1 and the instructions (*solution_len of them,
 * 0 if start already solves; solution_insns holds PCAGPU_MAX_KDEPTH), 0 if
 * there is none, or an error: PCAGPU_EINVAL also when max_depth -
 * start->depth exceeds PCAGPU_MAX_KDEPTH (the reference clamps at 16).  The
 * reference's 32-bit OEP table can prune a distinct state on a hash
 * collision or stop pruning when its probe window fills; exact keys do
 * neither, so results differ from it only where that happens. */
int     pcagpu_complete(pcagpu_t *h, const pca_search_state_t *start, int max_depth,
                        uint16_t *solution_insns, int *solution_len);
/* src/search_state.c:447-450 search_exhaustive_states_explored: children
 * generated by the last pcagpu_complete on h (every level's full child
 * list, so not the reference's DFS count, which stops at the first hit). */
uint64_t pcagpu_complete_states_explored(pcagpu_t *h);

/* ---- continuous-batching serving (the request loop above the path) ----
 * No reference counterpart: the reference solves one task per process.
 * Requests (task + depth budget) are queued; each pcagpu_server_step()
 * admits queued requests into resident slots while max_slots and the HBM
 * budget allow, stages the next split-depth prefix batch of every active
 * task, scans all of them in ONE persistent kernel launch, and resolves the
 * tasks whose depth is complete.  Results are identical to pcagpu_solve on
 * the task alone (split D-1).  Single-threaded: one caller drives step(). */
typedef struct pcagpu_server pcagpu_server_t;

#define PCAGPU_REQ_QUEUED  0
#define PCAGPU_REQ_ACTIVE  1
#define PCAGPU_REQ_DONE    2   /* optimum proven, or none up to the budget  */
#define PCAGPU_REQ_FAILED  3

typedef struct {
    int      device;
    uint32_t max_slots;          /* tasks resident at once (default 64)        */
    uint64_t hbm_budget;         /* device bytes the slots may hold (default 200 GiB) */
    uint32_t prefixes_per_item;  /* prefixes a task adds to one launch (default 16384) */
    uint32_t warm_slots;         /* slots opened by server_open, so the first
                                    requests pay no allocation (default 4;
                                    0xFFFFFFFF: none)                          */
    uint32_t step_prefixes;      /* prefixes one step scans over all tasks: a
                                    task adds max(prefixes_per_item,
                                    step_prefixes / active tasks) (default 262144) */
    uint32_t hit_capacity;       /* device hit records per slot before a batch
                                    is rescanned alone with a larger buffer
                                    (default 65536)                           */
    uint32_t fail_step;          /* testing: the n-th step (1-based) fails
                                    after staging, as a failed allocation
                                    would (0: off)                            */
} pcagpu_server_opts_t;

typedef struct {
    uint64_t       id;
    int            status;          /* PCAGPU_REQ_*                               */
    int            optimal_depth;   /* -1: no program up to max_depth             */
    long           num_solutions;
    int            max_depth;
    double         t_submit, t_admit, t_done;   /* seconds, server clock          */
    pca_solution_t first;           /* first program in canonical order           */
} pcagpu_request_t;

typedef struct {
    uint64_t submitted, admitted, completed, failed;
    uint64_t launches, items;       /* scan launches, task batches in them        */
    uint64_t kernel_ns, candidates, suffixes;
    uint64_t slots, slots_opened, device_bytes;
    /* host time of the steps, by phase: admission (re-tasking, the first
     * depth), staging the batches, waiting for the launches, and finished
     * depths (resolution, the next depth's prefixes) */
    uint64_t admit_ns, stage_ns, wait_ns, resolve_ns;
    uint64_t rescans;               /* batches rescanned alone after their hits
                                       overflowed the slot's buffer           */
} pcagpu_server_stats_t;

int    pcagpu_server_open(const pcagpu_server_opts_t *opts, pcagpu_server_t **out);
void   pcagpu_server_close(pcagpu_server_t *s);
int    pcagpu_server_submit(pcagpu_server_t *s, const pca_task_t *task, int max_depth, uint64_t *id);
/* One scheduling iteration; returns requests still queued or active. */
int    pcagpu_server_step(pcagpu_server_t *s);
/* Completed requests since the last poll (each reported once). */
int    pcagpu_server_poll(pcagpu_server_t *s, pcagpu_request_t *out, int cap);
/* Canonical solutions of a completed request: the first
 * min(cap, PCAGPU_SERVER_KEEP, total) are written; returns the total count
 * (which can exceed what is kept). */
#define PCAGPU_SERVER_KEEP 4096
long   pcagpu_server_solutions(pcagpu_server_t *s, uint64_t id, pca_solution_t *out, long cap);
int    pcagpu_server_get_stats(pcagpu_server_t *s, pcagpu_server_stats_t *st);
double pcagpu_server_clock(pcagpu_server_t *s);

#ifdef __cplusplus
}
#endif
#endif
