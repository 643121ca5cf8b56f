/*
 * pca_oracle.h -- CPU restatement of the PCA-16 reference solver.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for the parity tests, the
 * smoke check and the CPU-baseline leg of bench.py.  The product library
 * (pca_amd/) never links or calls it.
 *
 * What it restates (reference file:line):
 *   VM ............ src/vm.c:14-217, src/pca.h:24-139
 *   task loader ... tasks/spec.c:64-184 (quirks kept: max_cycles default
 *                   1000, 32-test cap, values & 0xFFFF, strstr key lookup)
 *   evaluator ..... tasks/spec.c:188-215
 *   disassembler .. src/asm.c:404-503
 *   grammar ....... tools/enumerate.c:196-267 (branch mode),
 *                   src/search_state.c:280-330 (branchless mode)
 *   IDDFS ......... tools/enumerate.c:271-461, :537-662
 *   search state .. src/search_state.c:15-239
 *   completion .... src/search_state.c:339-445 (search_exhaustive_complete)
 *
 * OEP modes:
 *   OR_OEP_EXACT  keys are the full per-test state bytes (regs r0..r7 lo/hi
 *                 and flags, the bytes the reference hashes); no collisions,
 *                 no saturation.  This is the semantics the GPU path matches.
 *   OR_OEP_REF32  the reference's own table: FNV-1a-32 (branch mode) or the
 *                 low 32 bits of FNV-1a-64 (branchless), 2^bits slots,
 *                 key = hash|1, 32 linear probes, full window => no prune
 *                 (tools/enumerate.c:43-77).  Used to pin this file
 *                 byte-for-byte against the reference binary.  An exact
 *                 shadow set counts the events where the 32-bit table
 *                 disagrees with exact keys (collision prunes, saturations).
 *   OR_OEP_OFF    no pruning (--no-oep).
 */
#ifndef PCA_ORACLE_H
#define PCA_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_NUM_REGS   8
#define OR_MAX_TESTS  32
#define OR_MAX_PORTS  8
#define OR_PROG_SIZE  256

/* Byte layout identical to task_spec_t (tasks/spec.h:20-34, 1128 B). */
typedef struct {
    uint16_t in[OR_MAX_PORTS];
    uint16_t out[OR_MAX_PORTS];
} or_test_t;

typedef struct {
    char      name[64];
    uint64_t  max_cycles;
    uint8_t   input_ports[OR_MAX_PORTS];
    int       num_inputs;
    uint8_t   output_ports[OR_MAX_PORTS];
    int       num_outputs;
    or_test_t tests[OR_MAX_TESTS];
    int       num_tests;
} or_task_t;

typedef struct or_vm {
    uint16_t r[OR_NUM_REGS];
    uint16_t mem[256];
    uint16_t prog[OR_PROG_SIZE];
    uint8_t  pc, sp;
    uint16_t stack[16];
    uint8_t  flags, status;
    uint64_t cycles;
    uint16_t ports[8];
    /* I/O hooks, as in pca_vm_t (src/pca.h:105-120); the solver never sets
     * them but the step checks them exactly as the reference does */
    void (*on_write)(struct or_vm *vm, uint8_t port, uint16_t val);
    void (*on_read)(struct or_vm *vm, uint8_t port);
} or_vm_t;

enum { OR_RUNNING = 0, OR_HALTED = 1, OR_TIMEOUT = 2, OR_FAULT = 3 };
enum { OR_OEP_OFF = 0, OR_OEP_EXACT = 1, OR_OEP_REF32 = 2 };

/* ---- VM / task ---- */
void or_vm_init(or_vm_t *vm);
void or_vm_load(or_vm_t *vm, const uint16_t *prog, size_t len);
void or_vm_step(or_vm_t *vm);
int  or_vm_run(or_vm_t *vm, uint64_t max_cycles);
/* One ALU evaluation with the VM's exact semantics: returns the value that
 * would be written and the flags after; *fault = 1 for DIV/MOD by zero. */
uint16_t or_alu(int op, uint16_t a, uint16_t b, uint8_t flags_in,
                uint8_t *flags_out, int *fault);

int  or_task_load(const char *path, or_task_t *task);
int  or_task_parse(const char *json, or_task_t *task);
int  or_task_evaluate(const or_task_t *task, const uint16_t *prog, int len);
void or_disasm(uint16_t insn, char *buf, size_t len);

/* ---- candidate grammars ---- */
int or_gen_candidates(uint16_t *out, int max, uint8_t live, int pos,
                      int kernel_depth, int use_branches);
int or_gen_candidates_branchless(uint16_t *out, int max, uint8_t live);
uint8_t or_live_after(uint8_t live, uint16_t insn);

/* ---- hashes (tools/enumerate.c:439-452, src/search_state.c:190-219) ---- */
uint32_t or_fnv32_bytes(uint32_t h, const uint8_t *p, size_t n);
uint64_t or_fnv64_bytes(uint64_t h, const uint8_t *p, size_t n);

/* ---- IDDFS solver ---- */
typedef struct {
    int    use_branches;    /* 1 = branch mode (reference default)          */
    int    oep_mode;        /* OR_OEP_*                                     */
    int    oep_bits;        /* REF32 table bits (reference: 20)             */
    int    max_solutions;   /* reference: 100; <= 0 means keep every one    */
    int    start_depth;     /* first kernel depth of the IDDFS              */
    int    max_depth;       /* last kernel depth                            */
    double time_limit;      /* seconds; 0 = none (stops the DFS mid-depth)  */
    int    print;           /* print reference-format stdout                */
    int    shadow_exact;    /* REF32: keep an exact shadow set to count the
                               events where the 32-bit table diverges       */
} or_opts_t;

typedef struct {
    int      optimal_depth;             /* -1 if none                        */
    int      num_solutions;             /* kept (capped) solutions           */
    uint16_t (*solutions)[16];          /* kernel programs incl. scaffold    */
    int      *solution_lens;
    uint64_t tried[17];                 /* per kernel depth                  */
    uint64_t leaf_tried[17];            /* leaf (program, out-reg) evals     */
    uint64_t pruned[17];                /* per kernel depth                  */
    uint64_t ref_collisions[17];        /* REF32: prunes of a distinct state */
    uint64_t ref_saturations[17];       /* REF32: full windows               */
    int      cap_overflow;              /* a leaf pushed past the cap        */
    int      timed_out;
    double   seconds;
    int      depths_done;               /* last completed depth              */
} or_result_t;

void or_opts_default(or_opts_t *o);
int  or_solve(const or_task_t *task, const or_opts_t *opts, or_result_t *res);
void or_result_free(or_result_t *res);

/* ---- slices used by the GPU parity tests ---- */

/* Exact-OEP branch-mode DFS of the depth-D tree down to split s: writes the
 * surviving length-s prefixes in DFS order (kernel insns, 8 slots each) and
 * their live masks.  Returns the count, or -1 if more than cap survive.
 * This is hybrid_enumerate's emission (gpu/hybrid.cu:524-588) with exact
 * keys and no early stop. */
long or_frontier(const or_task_t *task, int D, int s,
                 uint16_t *insns_out, uint8_t *live_out, long cap);

/* Brute-force suffix of length k = D - s below one prefix, with the
 * branch-mode grammar at every suffix position (liveness tracked linearly,
 * BR offsets 1..D-pos-1) and no OEP inside the suffix.  Writes every
 * solving (suffix-index path, out-reg) in canonical order as rows of
 * k+1 uint16: the k candidate-list indices then the out register.  Returns
 * the hit count (may exceed cap: only the first cap rows are written). */
long or_suffix_hits(const or_task_t *task, int D, const uint16_t *prefix,
                    int s, uint8_t live, uint16_t *rows, long cap,
                    uint64_t *candidates_out);

/* The last kernel position of depth D below n prefixes of D-1 instructions
 * (kernel insns, 8 slots each, and live masks), brute force on the VM with
 * nthreads threads: every child whose replay state has a register r1..r7
 * equal to the target on every test, in DFS order, as rows of 4 uint32:
 * prefix index, candidate index, that register mask, flags (bit 0: the
 * program cannot pass on some test whatever it outputs; bit 1: the first
 * child of the rows so far with this state, i.e. kept by the per-depth OEP
 * when the prefixes are the whole frontier).  Returns the row count (only
 * the first cap are written) or -1. */
long or_last_level(const or_task_t *task, int D, const uint16_t *insns, const uint8_t *live, long n,
                   int nthreads, uint32_t *rows, long cap, uint64_t *children_out);
/* Exhaustive ALU digest over a in [0,2^16), b in [b_lo,b_hi): see the .c. */
uint64_t or_alu_digest(int op, uint32_t b_lo, uint32_t b_hi);

/* Replay state of a partial kernel (scaffold + kernel + HLT, 100 cycles),
 * exactly as the reference OEP hashes it: per test r0..r7 then flags.
 * out must hold num_tests*9 uint16 (flags stored in the 9th). */
void or_replay_state(const or_task_t *task, const uint16_t *kernel, int len,
                     uint16_t *out);

/* ---- exhaustive leaf completion (src/search_state.c:339-445) ---- */

/* Byte layout identical to search_state_t (src/search_state.h:19-25, 546 B):
 * per-test registers and flags, live mask, kernel instructions so far. */
typedef struct {
    uint16_t regs[OR_MAX_TESTS][8];
    uint8_t  flags[OR_MAX_TESTS];
    uint8_t  live;
    uint8_t  depth;
} or_state_t;

/* search_state_init_root / search_state_step / search_state_leaf_outputs
 * (src/search_state.c:34-46, :49-187, :223-239). */
void    or_state_root(const or_task_t *task, or_state_t *st);
int     or_state_step(const or_task_t *task, const or_state_t *src, uint16_t insn, or_state_t *dst);
uint8_t or_state_leaf(const or_task_t *task, const or_state_t *st);

/* search_exhaustive_complete: IDDFS over 1..max_depth - start->depth
 * appended branchless instructions (at most 16), per-position OEP tables
 * cleared per iteration; returns 1 with the first completion in DFS order
 * (insns, *len; len 0 when the start state already solves), 0 if none, -1
 * on allocation failure.  oep_mode OR_OEP_REF32 is the reference's 2^20
 * table of 32-bit keys with 32 probes; OR_OEP_EXACT uses full state bytes.
 * *explored counts successful child steps (search_exhaustive_states_explored). */
int or_complete(const or_task_t *task, const or_state_t *start, int max_depth, int oep_mode,
                uint16_t *insns, int *len, uint64_t *explored);

#ifdef __cplusplus
}
#endif
#endif
